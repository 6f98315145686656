"""The CPU oracle's mesh set-up, pinned by hand-traced known answers on the
reference's dummy mesh and by invariants that any mistake in the stencil or
numbering would break."""
import json
import os

import numpy as np
import pytest

from ufemism_amd import geometry, meshgen

po = pytest.importorskip("oracle.pyoracle")

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dummy_mesh_AaAc.json")))


def rows(a, counts):
    return [a[i, :c].tolist() for i, c in enumerate(counts)]


def test_dummy_mesh_known_answers():
    m = meshgen.dummy_mesh()
    om = po.OracleMesh(m).build()
    assert om.nAc == GOLD["nAc"]
    assert rows(om["iAci"], m.nC) == GOLD["iAci"]
    assert om["Aci"].tolist() == GOLD["Aci"]
    assert om["edge_index_Ac"].tolist() == GOLD["edge_index_Ac"]
    assert om["nCAaAc"].tolist() == GOLD["nCAaAc"]
    assert rows(om["CAaAc"], om["nCAaAc"]) == GOLD["CAaAc"]
    assert om.colour_status == 0
    assert om["colour"].tolist() == GOLD["colour"]
    assert om.colour_nV == GOLD["colour_nV"]
    assert np.allclose(om["Nx"][4, :5], GOLD["centre_vertex_Nx"], atol=1e-15)
    assert np.allclose(om["Ny"][4, :5], GOLD["centre_vertex_Ny"], atol=1e-15)


@pytest.fixture(scope="module", params=[("dome", 2500, 7), ("marine", 4000, 8)])
def built(request):
    kind, n, seed = request.param
    res = geometry.dome_resolution() if kind == "dome" else geometry.marine_resolution()
    m = meshgen.generate_mesh(res, n, seed=seed)
    om = po.OracleMesh(m).build()
    assert om.colour_status == 0
    return m, om


def test_Ac_numbering_is_first_encounter(built):
    """The scan numbers (vi, ci) with C(vi,ci) > vi in scan order."""
    m, om = built
    expect = np.zeros_like(om["iAci"])
    k = 0
    for vi in range(m.nV):
        for ci in range(m.nC[vi]):
            vj = m.C[vi, ci] - 1
            if vj > vi:
                k += 1
                expect[vi, ci] = k
    for vi in range(m.nV):
        for ci in range(m.nC[vi]):
            vj = m.C[vi, ci] - 1
            if vj < vi:
                cj = list(m.C[vj, :m.nC[vj]] - 1).index(vi)
                expect[vi, ci] = expect[vj, cj]
    assert np.array_equal(om["iAci"], expect)
    Aci = om["Aci"] - 1
    for a in range(om.nAc):
        vi, vj = Aci[a, 0], Aci[a, 1]
        assert vi < vj and om["iAci"][vi, list(m.C[vi, :m.nC[vi]] - 1).index(vj)] == a + 1


def test_left_right_vertices(built):
    """(vi, vj, vl) and (vj, vi, vr) are counter-clockwise triangles of the
    mesh; a boundary segment has one triangle, whose third vertex is vl."""
    m, om = built
    tris = {tuple(np.roll(t, -int(np.argmin(t)))) for t in (m.Tri - 1).tolist()}
    sets = {frozenset(t) for t in (m.Tri - 1).tolist()}
    canon = lambda t: tuple(np.roll(t, -int(np.argmin(t))))
    Aci = om["Aci"] - 1
    for a in range(om.nAc):
        vi, vj, vl, vr = Aci[a]
        if om["edge_index_Ac"][a] == 0:
            assert canon([vi, vj, vl]) in tris
            assert canon([vj, vi, vr]) in tris
        else:
            assert frozenset([vi, vj, vl]) in sets and vr == 0


def test_AaAc_connectivity_symmetric_and_ccw(built):
    m, om = built
    nA, nV = om.nVAaAc, m.nV
    n = om["nCAaAc"]
    C = om["CAaAc"] - 1
    P = om["VAaAc"]
    nbr = [set(C[i, :n[i]].tolist()) for i in range(nA)]
    for i in range(nA):
        for j in nbr[i]:
            assert i in nbr[j]
    for i in range(nV, nA):
        assert n[i] == (4 if om["edge_index_Ac"][i - nV] > 0 else 6)
        ang = np.arctan2(P[C[i, :n[i]], 1] - P[i, 1], P[C[i, :n[i]], 0] - P[i, 0])
        turn = np.diff(np.unwrap(ang))
        assert np.all(turn > 0), i                      # counter-clockwise
    for vi in range(nV):
        assert np.array_equal(C[vi, :n[vi]], om["iAci"][vi, :m.nC[vi]] - 1 + nV)


def _fields(P):
    x, y = P[:, 0] / 1e5, P[:, 1] / 1e5
    return x, y


def test_first_derivative_weights_exact_on_linear_fields(built):
    m, om = built
    for pts, nc, C, Nx, Ny in ((m.V, m.nC, m.C, om["Nx"], om["Ny"]),
                              (om["VAaAc"], om["nCAaAc"], om["CAaAc"], om["Nx_AaAc"], om["Ny_AaAc"])):
        f = 3.0 * pts[:, 0] - 2.0 * pts[:, 1] + 7.0
        n = nc
        idx = np.arange(len(n))
        dx = Nx[idx, n] * f
        dy = Ny[idx, n] * f
        for c in range(int(n.max())):
            has = n > c
            dx[has] += Nx[idx[has], c] * f[C[idx[has], c] - 1]
            dy[has] += Ny[idx[has], c] * f[C[idx[has], c] - 1]
        assert np.allclose(dx, 3.0, rtol=0, atol=1e-9)
        assert np.allclose(dy, -2.0, rtol=0, atol=1e-9)


def test_second_derivative_weights_annihilate_linear_fields(built):
    m, om = built
    for pts, nc, C, Ns, edge in ((m.V, m.nC, m.C, (om["Nxx"], om["Nxy"], om["Nyy"]), m.edge_index > 0),):
        f = 3.0 * pts[:, 0] - 2.0 * pts[:, 1] + 7.0
        idx = np.arange(len(nc))
        for N in Ns:
            d = N[idx, nc] * f
            for c in range(int(nc.max())):
                has = nc > c
                d[has] += N[idx[has], c] * f[C[idx[has], c] - 1]
            ok = np.isfinite(d)
            assert np.all(ok | edge)
            scale = np.abs(N[idx, nc]).max() * np.abs(f).max()
            assert np.abs(d[ok]).max() <= 1e-9 * scale


def test_Ac_derivatives_exact_on_linear_fields(built):
    m, om = built
    f = 3.0 * m.V[:, 0] - 2.0 * m.V[:, 1] + 7.0
    n = om.nAc
    dx, dy, dp, do = (np.zeros(n) for _ in range(4))
    po.op("get_mesh_derivatives_Ac", om, f, dx, dy, dp, do)
    assert np.allclose(dx, 3.0, atol=1e-9) and np.allclose(dy, -2.0, atol=1e-9)
    A = om["Aci"] - 1
    d = m.V[A[:, 1]] - m.V[A[:, 0]]
    L = np.hypot(d[:, 0], d[:, 1])
    tp = (3.0 * d[:, 0] - 2.0 * d[:, 1]) / L         # derivative along vi -> vj
    to = (-2.0 * d[:, 0] - 3.0 * d[:, 1]) / L        # along the left normal
    assert np.allclose(dp, tp, atol=1e-9) and np.allclose(do, to, atol=1e-9)


def test_derivative_convergence_orders():
    """sin x cos y on refined regular meshes: first derivatives converge at
    second order and second derivatives at least at first order
    (documentation/UFEMISM_documentation.tex:1916-1931)."""
    errs1, errs2, hs = [], [], []
    L = 800e3
    k = np.pi / L
    for n in (2000, 8000, 32000):
        m = meshgen.generate_mesh(lambda x, y: np.ones_like(x), n, seed=3, jitter=0.0)
        om = po.OracleMesh(m).build(colour=False)
        x, y = m.V[:, 0], m.V[:, 1]
        f = np.sin(k * x) * np.cos(k * y)
        dx, dy = np.zeros(m.nV), np.zeros(m.nV)
        xx, xy, yy = np.zeros(m.nV), np.zeros(m.nV), np.zeros(m.nV)
        po.op("get_mesh_derivatives", om, f, dx, dy)
        po.op("get_mesh_curvatures", om, f, xx, xy, yy)
        inner = (m.edge_index == 0) & (np.abs(x) < 0.7 * L) & (np.abs(y) < 0.7 * L)
        e1 = np.abs(dx - k * np.cos(k * x) * np.cos(k * y))[inner].max() / k
        e2 = np.abs(xx + k * k * f)[inner].max() / (k * k)
        errs1.append(e1)
        errs2.append(e2)
        hs.append(np.sqrt(4 * L * L / m.nV))
    o1 = np.polyfit(np.log(hs), np.log(errs1), 1)[0]
    o2 = np.polyfit(np.log(hs), np.log(errs2), 1)[0]
    assert o1 > 1.8, o1       # second order
    assert o2 > 0.9, o2       # at least first order (second on a regular lattice)


def test_buggy_AaAc_curvature_is_reproduced(built):
    """The AaAc curvature weighs every slot by the home value
    (src/mesh_ArakawaC_module.f90:651-656): it is d(ai) times the weight sum,
    and on a quadratic it does not return the curvature."""
    m, om = built
    P = om["VAaAc"]
    f = (P[:, 0] / 1e5) ** 2
    n = om.nVAaAc
    xx, xy, yy = np.zeros(n), np.zeros(n), np.zeros(n)
    po.op("get_mesh_curvatures_AaAc", om, f, xx, xy, yy)
    idx = np.arange(n)
    nc = om["nCAaAc"]
    wsum = om["Nxx_AaAc"][idx, nc].copy()
    wabs = np.abs(wsum)
    for c in range(int(nc.max())):
        has = nc > c
        wsum[has] += om["Nxx_AaAc"][idx[has], c]
        wabs[has] += np.abs(om["Nxx_AaAc"][idx[has], c])
    ok = np.isfinite(wsum)
    # equal up to the rounding of the cancelling sum
    assert np.all(np.abs(xx[ok] - f[ok] * wsum[ok]) <= 1e-14 * np.abs(f[ok]) * wabs[ok])
    inner = ok & (np.abs(P[:, 0]) < 5e5) & (np.abs(P[:, 1]) < 5e5)
    assert np.abs(xx[inner]).max() < 1e-3 * 2e-10      # the true value is 2e-10 per m^2


def test_neumann_is_idempotent_and_preserves_constants(built):
    m, om = built
    rng = np.random.default_rng(0)
    for name, n in (("apply_Neumann_boundary", m.nV), ("apply_Neumann_boundary_AaAc", om.nVAaAc)):
        d = rng.standard_normal(n)
        po.op(name, om, d)
        d2 = d.copy()
        po.op(name, om, d2)
        assert np.array_equal(d, d2)
        c = np.full(n, 4.25)
        po.op(name, om, c)
        assert np.allclose(c, 4.25, rtol=1e-15)


def test_colouring_valid_and_ascending(built):
    m, om = built
    col = om["colour"]
    C = om["CAaAc"] - 1
    for i in range(om.nVAaAc):
        nb = C[i, :om["nCAaAc"][i]]
        assert 1 <= col[i] <= 5 and np.all(col[nb] != col[i])
    for c in range(5):
        lst = om["colour_vi"][:om.colour_nV[c], c]
        assert np.all(np.diff(lst) > 0) and np.all(col[lst - 1] == c + 1)
    assert sum(om.colour_nV) == om.nVAaAc


def test_colouring_identify_path_aborts():
    """A planar graph of minimum degree five (the icosahedron) forces the
    identify() branch, which the reference aborts on (src/mesh_five_colour_module.f90:452)."""
    ico = [[2, 3, 4, 5, 6], [1, 3, 6, 7, 8], [1, 2, 4, 8, 9], [1, 3, 5, 9, 10], [1, 4, 6, 10, 11],
           [1, 2, 5, 7, 11], [2, 6, 8, 11, 12], [2, 3, 7, 9, 12], [3, 4, 8, 10, 12],
           [4, 5, 9, 11, 12], [5, 6, 7, 10, 12], [7, 8, 9, 10, 11]]
    status, _ = po.colour_graph(ico)
    assert status == 4


def test_colouring_rotation_path_gets_stuck():
    """The Q5 branch looks for two non-adjacent neighbours x, y of v with
    deg < C%nconmax; without such a pair it returns v to the rear of Q5
    (src/mesh_five_colour_module.f90:222-228).  On the icosahedron every vertex
    has degree 5, so with nC_mem = 5 no pair qualifies: Q5 rotates until the
    reduction gives up after 2 nVAaAc iterations (:167-170, the 'reduce got
    stuck' abort, status 5).  With nC_mem = 6 the first vertex has a pair and
    identify() aborts (status 4); a threshold of deg <= k, or k taken from
    anything but nC_mem, would change one of the two."""
    ico = [[2, 3, 4, 5, 6], [1, 3, 6, 7, 8], [1, 2, 4, 8, 9], [1, 3, 5, 9, 10], [1, 4, 6, 10, 11],
           [1, 2, 5, 7, 11], [2, 6, 8, 11, 12], [2, 3, 7, 9, 12], [3, 4, 8, 10, 12],
           [4, 5, 9, 11, 12], [5, 6, 7, 10, 12], [7, 8, 9, 10, 11]]
    assert po.colour_graph(ico, nC_mem=5)[0] == 5
    assert po.colour_graph(ico, nC_mem=6)[0] == 4


def test_colouring_small_known_graph():
    """A wheel with five spokes: the hub has degree 5, the rim degree 3; the
    LIFO reduction deletes rim vertices from the top of Q4 until five remain."""
    wheel = [[2, 3, 4, 5, 6], [1, 3, 6], [1, 2, 4], [1, 3, 5], [1, 4, 6], [1, 5, 2]]
    status, col = po.colour_graph(wheel)
    assert status == 0
    # Q4 starts [2,3,4,5,6], Q5 [1]; deleting 6 drops the hub to degree 4
    # (Q4 becomes [2,3,4,5,1]); five vertices remain and are coloured 1..5 in
    # Q4 order (2, 3, 4, 5, 1 -> colours 1..5); 6 then takes the smallest
    # colour unused by its saved neighbours 1, 5, 2 (colours 5, 4, 1): 2.
    assert col.tolist() == [5, 1, 2, 3, 4, 2]
