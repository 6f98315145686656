"""The oracle's primary-mesh geometry (src/mesh_help_functions_module.f90):
circumcentres, Voronoi cell areas with the boundary cropping, connection
widths, triangle areas and resolution.  Pinned by closed forms on the
reference's dummy mesh, by qhull's independent Voronoi diagram, by a raster
count of nearest vertices, by exact tiling of the domain and by the
reference's own failure path for a degenerate triangle."""
import numpy as np
import pytest

from ufemism_amd import geometry, meshgen

po = pytest.importorskip("oracle.pyoracle")


def test_dummy_mesh_closed_forms():
    """Four right triangles around the centre: their circumcentres are the
    edge midpoints, the centre's cell is the diamond through them (half the
    square), each corner keeps an eighth, centre-corner widths are the
    midpoint distance sqrt(2) and the boundary connections have width 0."""
    m = meshgen.dummy_mesh()
    g = po.mesh_geometry(m)
    assert g["invalid_cc"] == 0 and g["rotated"] == 0 and g["outside"] == 0 and g["cw_fails"] == 0
    assert np.array_equal(g["Tricc"], np.array([[0.0, -1.0], [1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]]))
    assert np.array_equal(g["A"], np.array([0.5, 0.5, 0.5, 0.5, 2.0]))
    assert np.array_equal(g["TriA"], np.ones(4))
    r2 = np.sqrt(2.0)
    for vi in range(4):
        row = g["Cw"][vi, :3]
        assert np.allclose(row, [0.0, r2, 0.0], rtol=1e-15), (vi, row)   # C(vi) = next corner, centre, previous corner
    assert np.allclose(g["Cw"][4, :4], r2, rtol=1e-15)
    assert np.allclose(g["R"], r2, rtol=1e-15)
    assert g["resolution_min"] == g["resolution_max"] == g["R"][0]


@pytest.fixture(scope="module", params=["dome2k", "marine5k"])
def mesh_geo(request):
    m = {"dome2k": lambda: geometry.dome(2000, seed=3).mesh,
         "marine5k": lambda: geometry.marine(5000, seed=8).mesh}[request.param]()
    return m, po.mesh_geometry(m)


def test_circumcentres_are_equidistant(mesh_geo):
    m, g = mesh_geo
    assert g["invalid_cc"] == 0
    P = m.V[g["Tri"] - 1]                                 # (nTri, 3, 2)
    d = np.hypot(*(P - g["Tricc"][:, None, :]).transpose(2, 0, 1))
    assert np.allclose(d[:, 1], d[:, 0], rtol=1e-8) and np.allclose(d[:, 2], d[:, 0], rtol=1e-8)
    # the same triangles as the mesh, vertex sets unchanged
    assert np.array_equal(np.sort(g["Tri"], axis=1), np.sort(m.Tri, axis=1))


def test_areas_tile_the_domain(mesh_geo):
    m, g = mesh_geo
    area = (m.xmax - m.xmin) * (m.ymax - m.ymin)
    assert g["outside"] == 0
    assert abs(g["A"].sum() / area - 1.0) < 1e-12
    assert abs(g["TriA"].sum() / area - 1.0) < 1e-12
    assert np.all(g["A"] > 0) and np.all(g["TriA"] > 0)


def _cropped_cells(m, g):
    cc = g["Tricc"]
    out = (cc[:, 0] < m.xmin) | (cc[:, 0] > m.xmax) | (cc[:, 1] < m.ymin) | (cc[:, 1] > m.ymax)
    bad_tris = np.nonzero(out)[0] + 1
    cells = set()
    for vi in range(m.nV):
        if np.isin(m.iTri[vi, : m.niTri[vi]], bad_tris).any():
            cells.add(vi)
    return out, sorted(cells)


def test_voronoi_areas_match_a_raster_count(mesh_geo):
    """Every cell's area against the count of raster points whose nearest
    vertex it is (the Voronoi cell clipped to the domain, independent of the
    reference's polygon construction); cells touching a circumcentre outside
    the domain exercise the cropping."""
    from scipy.spatial import cKDTree
    m, g = mesh_geo
    out, cropped = _cropped_cells(m, g)
    assert out.sum() > 0 and len(cropped) > 0
    n = 2000
    h = (m.xmax - m.xmin) / n
    x = m.xmin + h * (np.arange(n) + 0.5)
    X, Y = np.meshgrid(x, x)
    _, owner = cKDTree(m.V).query(np.stack([X.ravel(), Y.ravel()], 1))
    raster = np.bincount(owner, minlength=m.nV) * h * h
    # a raster misplaces at most the pixels along the cell boundary: perimeter * h / 2
    tol = 4.0 * np.sqrt(g["A"]) * h / 2
    err = np.abs(raster - g["A"])
    assert np.all(err <= tol), (np.max(err / tol), np.argmax(err / tol))
    # the generator's kite estimate (no cropping) misses the cropped cells by far more
    kite_err = np.abs(m.A - raster)[cropped]
    assert np.max(kite_err / tol[cropped]) > 2


def test_voronoi_cells_stay_inside_the_domain(mesh_geo):
    m, g = mesh_geo
    _, cropped = _cropped_cells(m, g)
    for vi in cropped[:20]:
        Vor, outside = po.voronoi_cell_vertices(m, g["Tricc"], vi + 1)
        assert outside == 0
        assert np.all((Vor[:, 0] >= m.xmin) & (Vor[:, 0] <= m.xmax))
        assert np.all((Vor[:, 1] >= m.ymin) & (Vor[:, 1] <= m.ymax))


def test_connection_widths_against_qhull(mesh_geo):
    """Interior connections: the Voronoi ridge between two vertices, whose ends
    qhull computes independently; all connections: symmetric, and zero or the
    distance to the boundary where only one triangle is shared."""
    from scipy.spatial import Voronoi
    m, g = mesh_geo
    Cw = g["Cw"]
    assert g["cw_fails"] == 0
    vor = Voronoi(m.V)
    ridge = {}
    for (a, b), rv in zip(vor.ridge_points, vor.ridge_vertices):
        if -1 not in rv:
            p, q = vor.vertices[rv[0]], vor.vertices[rv[1]]
            ridge[(min(a, b), max(a, b))] = np.hypot(*(p - q))
    checked = 0
    for vi in range(m.nV):
        for ci in range(m.nC[vi]):
            vj = m.C[vi, ci] - 1
            back = list(m.C[vj, : m.nC[vj]] - 1).index(vi)
            assert Cw[vi, ci] == Cw[vj, back]
            if m.edge_index[vi] == 0 and m.edge_index[vj] == 0 and (min(vi, vj), max(vi, vj)) in ridge:
                w = ridge[(min(vi, vj), max(vi, vj))]
                assert abs(Cw[vi, ci] - w) <= 1e-6 * max(w, 1.0), (vi, vj, Cw[vi, ci], w)
                checked += 1
        assert np.all(Cw[vi, m.nC[vi]:] == 0.0)
    assert checked > m.nV
    # free vertices whose cell is not cropped: the fan area equals the kite sum
    # over the cell edges (distance x width / 4)
    _, cropped = _cropped_cells(m, g)
    free = [vi for vi in range(m.nV) if m.edge_index[vi] == 0 and vi not in set(cropped)]
    for vi in free[:300]:
        vj = m.C[vi, : m.nC[vi]] - 1
        dist = np.hypot(*(m.V[vj] - m.V[vi]).T)
        assert np.isclose(np.sum(dist * Cw[vi, : m.nC[vi]]) / 4.0, g["A"][vi], rtol=1e-12)


def test_resolution_is_the_shortest_connection(mesh_geo):
    m, g = mesh_geo
    for vi in range(0, m.nV, 7):
        vj = m.C[vi, : m.nC[vi]] - 1
        assert np.isclose(g["R"][vi], np.hypot(*(m.V[vj] - m.V[vi]).T).min(), rtol=1e-15)
    assert g["resolution_min"] == g["R"].min() and g["resolution_max"] == g["R"].max()


def test_degenerate_triangle_is_reported():
    """A collinear triangle has no circumcentre: the reference retries with the
    vertices rotated twice and then reports it (:867-895)."""
    m = meshgen.dummy_mesh()
    m.V = np.asfortranarray(m.V.copy())
    m.V[4] = [0.0, -1.0]                   # centre onto the south edge: triangle 1 collinear
    g = po.mesh_geometry(m)
    assert g["invalid_cc"] == 1 and g["rotated"] == 2
    assert np.all(g["Tricc"][0] == np.float32(1e30))
    # two rotations of (1, 2, 5): the vertex set and the orientation are kept
    assert list(g["Tri"][0]) == [5, 1, 2]
    assert np.array_equal(g["Tri"][1:], np.asarray(m.Tri)[1:])
    # compute_Tricc=False takes the given circumcentres
    m2 = meshgen.dummy_mesh()
    g2 = po.mesh_geometry(m2, compute_Tricc=False)
    assert np.array_equal(g2["Tricc"], np.asarray(m2.Tricc))


def test_generator_splits_encroached_boundary_segments():
    """A jittered marine mesh whose Delaunay triangulation has a triangle with
    its circumcentre outside the domain beyond a side it does not border: the
    reference cannot crop that Voronoi vertex and stops.  The generator splits
    the encroached boundary segment (one extra boundary vertex) and the mesh
    passes; meshes without such triangles are not touched."""
    res = geometry.marine_resolution()
    raw = meshgen.generate_mesh(res, 20213, seed=11062, split_encroached=False)
    assert po.mesh_geometry(raw)["outside"] > 0
    fixed = meshgen.generate_mesh(res, 20213, seed=11062)
    g = po.mesh_geometry(fixed)
    assert g["outside"] == 0 and g["cw_fails"] == 0
    assert fixed.nV == raw.nV + 1 and fixed.edge_index[-1] in (1, 3, 5, 7)
    m = geometry.dome(2000, seed=3).mesh
    same = meshgen.generate_mesh(geometry.dome_resolution(), 2000, seed=3, split_encroached=False)
    assert np.array_equal(m.V, same.V)
