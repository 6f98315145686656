"""Primary-mesh conventions of the seeded generator (reference layout)."""
import numpy as np
import pytest

from ufemism_amd import geometry, meshgen


def test_dummy_mesh_is_the_reference_table():
    # src/mesh_creation_module.f90:1898-2015
    m = meshgen.dummy_mesh()
    assert m.V.tolist() == [[-1, -1], [1, -1], [1, 1], [-1, 1], [0, 0]]
    assert m.nC.tolist() == [3, 3, 3, 3, 4]
    assert m.C[:, :4].tolist() == [[2, 5, 4, 0], [3, 5, 1, 0], [4, 5, 2, 0], [1, 5, 3, 0], [1, 2, 3, 4]]
    assert m.iTri[:, :4].tolist() == [[1, 4, 0, 0], [2, 1, 0, 0], [3, 2, 0, 0], [4, 3, 0, 0], [1, 2, 3, 4]]
    assert m.Tri.tolist() == [[1, 2, 5], [2, 3, 5], [3, 4, 5], [4, 1, 5]]
    assert m.TriC.tolist() == [[2, 4, 0], [3, 1, 0], [4, 2, 0], [1, 3, 0]]
    assert m.edge_index.tolist() == [6, 4, 2, 8, 0]
    assert m.Tri_edge_index.tolist() == [5, 3, 1, 7]
    # Voronoi cells: corners are quarter squares of side 1, the centre the rest
    assert np.allclose(m.A, [0.5, 0.5, 0.5, 0.5, 2.0])


@pytest.fixture(scope="module", params=[("dome", 3000, 1), ("marine", 5000, 4)])
def mesh(request):
    kind, n, seed = request.param
    res = geometry.dome_resolution() if kind == "dome" else geometry.marine_resolution()
    return meshgen.generate_mesh(res, n, seed=seed)


def _cross(o, a, b):
    return (a[..., 0] - o[..., 0]) * (b[..., 1] - o[..., 1]) - (a[..., 1] - o[..., 1]) * (b[..., 0] - o[..., 0])


def test_corners_and_edges(mesh):
    V, e = mesh.V, mesh.edge_index
    assert np.array_equal(e[:4], [6, 4, 2, 8])
    assert V[:4].tolist() == [[mesh.xmin, mesh.ymin], [mesh.xmax, mesh.ymin], [mesh.xmax, mesh.ymax],
                              [mesh.xmin, mesh.ymax]]
    on = {1: V[:, 1] == mesh.ymax, 3: V[:, 0] == mesh.xmax, 5: V[:, 1] == mesh.ymin, 7: V[:, 0] == mesh.xmin}
    for side, mask in on.items():
        assert np.all(e[4:][mask[4:]] == side)
    assert np.all(e[4:][~(on[1] | on[3] | on[5] | on[7])[4:]] == 0)


def test_triangles_ccw_and_euler(mesh):
    P, T = mesh.V, mesh.Tri - 1
    assert np.all(_cross(P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]) > 0)
    nB = int(np.sum(mesh.edge_index > 0))
    assert mesh.nTri == 2 * mesh.nV - nB - 2          # Euler, triangulated polygon
    assert int(mesh.nC.sum()) // 2 == mesh.nV + mesh.nTri - 1


def test_fans_ccw_and_consistent(mesh):
    P = mesh.V
    for vi in range(mesh.nV):
        n, nt = mesh.nC[vi], mesh.niTri[vi]
        edge = mesh.edge_index[vi] > 0
        assert nt == (n - 1 if edge else n)
        C = mesh.C[vi, :n] - 1
        for k in range(nt):
            a, b = C[k], C[(k + 1) % n]
            t = mesh.Tri[mesh.iTri[vi, k] - 1] - 1
            assert set(t.tolist()) == {vi, a, b}
            assert _cross(P[vi], P[a], P[b]) > 0
        if edge:
            assert mesh.edge_index[C[0]] > 0 and mesh.edge_index[C[-1]] > 0


def test_TriC_neighbours(mesh):
    T, TC = mesh.Tri - 1, mesh.TriC - 1
    for ti in range(mesh.nTri):
        for n in range(3):
            tj = TC[ti, n]
            shared = set(T[ti].tolist()) - {T[ti, n]}
            if tj < 0:
                continue
            assert shared <= set(T[tj].tolist())
            assert ti in TC[tj].tolist()


def test_voronoi_areas_tile_the_domain(mesh):
    area = (mesh.xmax - mesh.xmin) * (mesh.ymax - mesh.ymin)
    assert abs(mesh.A.sum() / area - 1) < 2e-3
    assert np.all(mesh.A > 0)
