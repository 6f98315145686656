"""Seeded triangular-mesh generator in the reference's primary-mesh layout.

The reference builds its meshes by Ruppert refinement plus Lloyd smoothing,
merged across MPI ranks (src/mesh_creation_module.f90:354-485, 1344-1741);
that machinery is outside the ice-dynamics path.  This module writes its own
seeded generator instead and emits the *primary* mesh arrays exactly in the
form the reference stores them (src/data_types_module.f90:216-345):

* ``V``          (nV, 2)      vertex x, y [m]
* ``nC``, ``C``  (nV), (nV, nC_mem)  neighbours, counter-clockwise; for edge
  vertices the list runs from one boundary neighbour to the other
  (src/mesh_creation_module.f90:1898-1963 shows the convention on the
  five-vertex dummy mesh)
* ``niTri``, ``iTri``  triangles around a vertex, counter-clockwise, with
  ``iTri(k)`` the triangle between ``C(k)`` and ``C(k+1)``
* ``edge_index`` 0 interior, 1 N, 2 NE, 3 E, 4 SE, 5 S, 6 SW, 7 W, 8 NW;
  vertices 1..4 are the SW, SE, NE, NW corners
* ``Tri``, ``TriC``, ``Tricc``, ``Tri_edge_index`` triangle arrays
* ``A`` (Voronoi cell areas) and ``Cw`` (connection widths)

All index arrays are 1-based with 0 in unused slots and all 2-D arrays are
Fortran-ordered, so a pointer to them is exactly what a Fortran caller holds.

Point placement uses nested, jittered triangular lattices: level ``L`` has
spacing ``h0 / 2**L`` and contributes its new points wherever the target
resolution asks for that spacing.  Vertices after the four corners are sorted
along a Hilbert curve so that neighbouring vertices (and therefore the Ac and
AaAc points numbered from them) sit close together in memory.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NC_MEM = 16  # reference default nconmax (src/configuration_module.f90:78)


@dataclasses.dataclass
class Mesh:
    """Primary mesh in reference layout (Fortran-ordered, 1-based indices)."""

    xmin: float
    xmax: float
    ymin: float
    ymax: float
    nC_mem: int
    V: np.ndarray
    nC: np.ndarray
    C: np.ndarray
    niTri: np.ndarray
    iTri: np.ndarray
    edge_index: np.ndarray
    Tri: np.ndarray
    TriC: np.ndarray
    Tricc: np.ndarray
    Tri_edge_index: np.ndarray
    A: np.ndarray
    Cw: np.ndarray

    @property
    def nV(self) -> int:
        return int(self.V.shape[0])

    @property
    def nTri(self) -> int:
        return int(self.Tri.shape[0])

    @property
    def tol_dist(self) -> float:
        # src/mesh_creation_module.f90:1937 (tol = 1e-9)
        return ((self.xmax - self.xmin) + (self.ymax - self.ymin)) * 1e-9 / 2.0


def _f32(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


# --------------------------------------------------------------------------
# Five-vertex dummy mesh (src/mesh_creation_module.f90:1898-2015)
# --------------------------------------------------------------------------

def dummy_mesh(xmin=-1.0, xmax=1.0, ymin=-1.0, ymax=1.0, nC_mem=NC_MEM) -> Mesh:
    """The reference's five-vertex, four-triangle start mesh, table for table."""
    V = [[xmin, ymin], [xmax, ymin], [xmax, ymax], [xmin, ymax],
         [(xmin + xmax) / 2, (ymin + ymax) / 2]]
    nC = [3, 3, 3, 3, 4]
    Cl = [[2, 5, 4], [3, 5, 1], [4, 5, 2], [1, 5, 3], [1, 2, 3, 4]]
    niTri = [2, 2, 2, 2, 4]
    iTl = [[1, 4], [2, 1], [3, 2], [4, 3], [1, 2, 3, 4]]
    Tri = [[1, 2, 5], [2, 3, 5], [3, 4, 5], [4, 1, 5]]
    TriC = [[2, 4, 0], [3, 1, 0], [4, 2, 0], [1, 3, 0]]
    C = np.zeros((5, nC_mem), np.int32, order="F")
    iTri = np.zeros((5, nC_mem), np.int32, order="F")
    for i in range(5):
        C[i, : len(Cl[i])] = Cl[i]
        iTri[i, : len(iTl[i])] = iTl[i]
    mesh = Mesh(xmin, xmax, ymin, ymax, nC_mem,
                _f32(V, np.float64), _f32(nC, np.int32), C,
                _f32(niTri, np.int32), iTri, _f32([6, 4, 2, 8, 0], np.int32),
                _f32(Tri, np.int32), _f32(TriC, np.int32),
                np.zeros((4, 2), np.float64, order="F"),
                _f32([5, 3, 1, 7], np.int32),
                np.zeros(5, np.float64), np.zeros((5, nC_mem), np.float64, order="F"))
    _finish_geometry(mesh)
    return mesh


# --------------------------------------------------------------------------
# Point placement
# --------------------------------------------------------------------------

def _hilbert_index(x, y, order=16):
    """Hilbert-curve index of integer grid points (vectorised)."""
    x = x.astype(np.int64).copy()
    y = y.astype(np.int64).copy()
    d = np.zeros_like(x)
    s = 1 << (order - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant
        flip = ry == 0
        swap_x = flip & (rx == 1)
        x = np.where(swap_x, s - 1 - x, x)
        y = np.where(swap_x, s - 1 - y, y)
        xt = np.where(flip, y, x)
        y = np.where(flip, x, y)
        x = xt
        s >>= 1
    return d


def _lattice(h, xmin, xmax, ymin, ymax):
    dy = h * math.sqrt(3.0) / 2.0
    nj = int(math.ceil((ymax - ymin) / dy)) + 2
    ni = int(math.ceil((xmax - xmin) / h)) + 2
    j = np.arange(-1, nj)
    i = np.arange(-1, ni)
    J, I = np.meshgrid(j, i, indexing="ij")
    X = xmin + I * h + (J % 2) * (h / 2.0)
    Y = ymin + J * dy
    return X.ravel(), Y.ravel(), I.ravel(), J.ravel()


def _place_points(res, xmin, xmax, ymin, ymax, scale, rng, jitter, kappa=0.75):
    """Nested jittered triangular lattices following ``res(x, y) * scale``."""
    # Probe the resolution field to find the coarsest/finest spacing.
    gx, gy = np.meshgrid(np.linspace(xmin, xmax, 257), np.linspace(ymin, ymax, 257))
    rvals = res(gx.ravel(), gy.ravel()) * scale
    rmin, rmax = float(rvals.min()), float(rvals.max())
    h0 = rmax
    nlev = max(0, int(math.ceil(math.log2(h0 / (rmin * 0.999))))) + 1
    pts_x, pts_y = [], []
    for L in range(nlev):
        h = h0 / (2 ** L)
        X, Y, I, J = _lattice(h, xmin, xmax, ymin, ymax)
        if L > 0:
            # keep only the points that are new at this level: a point of
            # level L is in level L-1 iff its row index is even and its column
            # maps onto the coarse column (x = i*h + (j%2)*h/2 here, while the
            # coarse lattice has rows 2j' and x = i'*2h + (j'%2)*h)
            jc = J // 2
            coarse = ((J % 2) == 0) & (((I - (jc % 2)) % 2) == 0)
            new = ~coarse
            X, Y = X[new], Y[new]
            r = res(X, Y) * scale
            keep = r < h * 2.0 * kappa
            X, Y = X[keep], Y[keep]
        pts_x.append(X)
        pts_y.append(Y)
    X = np.concatenate(pts_x)
    Y = np.concatenate(pts_y)
    r = res(X, Y) * scale
    lstar = np.clip(np.floor(np.log2(2.0 * kappa * h0 / r)), 0, nlev - 1)
    hloc = h0 / np.power(2.0, lstar)
    X = X + rng.uniform(-jitter, jitter, X.shape) * hloc
    Y = Y + rng.uniform(-jitter, jitter, Y.shape) * hloc
    dist = np.minimum.reduce([X - xmin, xmax - X, Y - ymin, ymax - Y])
    inside = dist > 0.5 * hloc
    X, Y = X[inside], Y[inside]

    def hloc_at(x, y):
        rr = float(res(np.array([x]), np.array([y]))[0]) * scale
        ls = min(max(math.floor(math.log2(2.0 * kappa * h0 / rr)), 0), nlev - 1)
        return h0 / 2 ** ls

    def side(p0, p1):
        length = math.hypot(p1[0] - p0[0], p1[1] - p0[1])
        ts = [0.0]
        while True:
            x = p0[0] + (p1[0] - p0[0]) * ts[-1] / length
            y = p0[1] + (p1[1] - p0[1]) * ts[-1] / length
            step = hloc_at(x, y)
            if ts[-1] + step >= length - 0.5 * step:
                break
            ts.append(ts[-1] + step)
        ts.append(length)
        t = np.array(ts)
        # stretch so that the spacing stays smooth up to the far corner
        t = t * (length / t[-1])
        t = t[1:-1]
        return (p0[0] + (p1[0] - p0[0]) * t / length,
                p0[1] + (p1[1] - p0[1]) * t / length)

    corners = [(xmin, ymin), (xmax, ymin), (xmax, ymax), (xmin, ymax)]
    bx, by, bedge = [], [], []
    for (p0, p1, e) in [(corners[0], corners[1], 5), (corners[1], corners[2], 3),
                        (corners[3], corners[2], 1), (corners[0], corners[3], 7)]:
        sx, sy = side(p0, p1)
        # snap exactly onto the boundary line
        if e in (5, 1):
            sy = np.full_like(sx, p0[1])
        else:
            sx = np.full_like(sy, p0[0])
        bx.append(sx)
        by.append(sy)
        bedge.append(np.full(sx.shape, e, np.int32))
    return X, Y, np.concatenate(bx), np.concatenate(by), np.concatenate(bedge)


# --------------------------------------------------------------------------
# Triangulation -> reference arrays
# --------------------------------------------------------------------------

_START_ANGLE = {0: None, 5: 0.0, 6: 0.0, 3: 0.5 * math.pi, 4: 0.5 * math.pi,
                1: math.pi, 2: math.pi, 7: 1.5 * math.pi, 8: 1.5 * math.pi}


def _build_topology(P, edge_index, tris, xmin, xmax, ymin, ymax, nC_mem):
    nV = P.shape[0]
    # orient triangles counter-clockwise
    a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
    cross = ((P[b, 0] - P[a, 0]) * (P[c, 1] - P[a, 1])
             - (P[b, 1] - P[a, 1]) * (P[c, 0] - P[a, 0]))
    if np.any(cross == 0):
        raise ValueError("degenerate triangle in triangulation")
    cw = cross < 0
    tris = tris.copy()
    tris[cw, 1], tris[cw, 2] = tris[cw, 2].copy(), tris[cw, 1].copy()
    # spatial triangle order (Hilbert order of the centroid)
    cen = P[tris].mean(axis=1)
    span = max(xmax - xmin, ymax - ymin)
    gx = np.clip(((cen[:, 0] - xmin) / span * 65535), 0, 65535).astype(np.int64)
    gy = np.clip(((cen[:, 1] - ymin) / span * 65535), 0, 65535).astype(np.int64)
    tris = tris[np.argsort(_hilbert_index(gx, gy), kind="stable")]
    nTri = tris.shape[0]

    # wedges: for triangle (p,q,r) ccw, vertex p sees neighbour q then r ccw
    wv = tris.ravel()                                    # vertex
    wf = tris[:, [1, 2, 0]].ravel()                      # first neighbour
    ws = tris[:, [2, 0, 1]].ravel()                      # second neighbour
    wt = np.repeat(np.arange(nTri), 3)                   # triangle
    ang = np.arctan2(P[wf, 1] - P[wv, 1], P[wf, 0] - P[wv, 0]) % (2 * math.pi)
    start = np.zeros(nV)
    for e, s in _START_ANGLE.items():
        if s is not None:
            start[edge_index == e] = s
    interior = edge_index == 0
    rel = (ang - start[wv] + 1e-9) % (2 * math.pi)
    # interior vertices: start at the smallest absolute angle
    rel = np.where(interior[wv], ang, rel)
    order = np.lexsort((rel, wv))
    wv, wf, ws, wt = wv[order], wf[order], ws[order], wt[order]
    niTri = np.bincount(wv, minlength=nV).astype(np.int32)
    is_edge = ~interior
    nC = niTri + is_edge.astype(np.int32)
    if nC.max() > nC_mem:
        raise ValueError(f"vertex degree {nC.max()} exceeds nC_mem={nC_mem}")
    first = np.concatenate([[0], np.cumsum(niTri)[:-1]])
    pos = np.arange(wv.size) - first[wv]
    # consistency: consecutive wedges chain (second of k == first of k+1)
    nxt_ok = np.ones(wv.size, bool)
    same = wv[1:] == wv[:-1]
    nxt_ok[:-1] = ~same | (ws[:-1] == wf[1:])
    if not nxt_ok.all():
        raise ValueError("neighbour ordering around a vertex is not a chain")
    C = np.zeros((nV, nC_mem), np.int32, order="F")
    iTri = np.zeros((nV, nC_mem), np.int32, order="F")
    C[wv, pos] = wf + 1
    iTri[wv, pos] = wt + 1
    last = first + niTri - 1
    ev = np.nonzero(is_edge)[0]
    C[ev, niTri[ev]] = ws[last[ev]] + 1
    # interior chains must close
    iv = np.nonzero(interior)[0]
    if not np.all(ws[last[iv]] == wf[first[iv]]):
        raise ValueError("interior vertex fan does not close")

    # TriC: triangle across the edge opposite vertex n
    ea = tris[:, [1, 2, 0]].ravel().astype(np.int64)
    eb = tris[:, [2, 0, 1]].ravel().astype(np.int64)
    key = ea * nV + eb
    rkey = eb * nV + ea
    srt = np.argsort(key)
    idx = np.searchsorted(key[srt], rkey)
    idx = np.clip(idx, 0, key.size - 1)
    found = key[srt][idx] == rkey
    TriC = np.where(found, srt[idx] // 3 + 1, 0).reshape(nTri, 3).astype(np.int32)
    return tris, nC, C, niTri, iTri, TriC


def _circumcentres(P, tris):
    a, b, c = P[tris[:, 0]], P[tris[:, 1]], P[tris[:, 2]]
    bx, by = b[:, 0] - a[:, 0], b[:, 1] - a[:, 1]
    cx, cy = c[:, 0] - a[:, 0], c[:, 1] - a[:, 1]
    d = 2.0 * (bx * cy - by * cx)
    b2, c2 = bx * bx + by * by, cx * cx + cy * cy
    ux = (cy * b2 - by * c2) / d
    uy = (bx * c2 - cx * b2) / d
    return np.stack([a[:, 0] + ux, a[:, 1] + uy], axis=1)


def _tri_edge_index(mesh: Mesh):
    """Side index of boundary triangles, following the boundary walk of
    src/mesh_help_functions_module.f90:1248-1322 (W, N, E, S, then the
    last triangle of each side is corrected)."""
    T = np.zeros(mesh.nTri, np.int32)
    C, nC, iTri, niTri, V = mesh.C, mesh.nC, mesh.iTri, mesh.niTri, mesh.V
    tol = mesh.tol_dist

    def step(vi_prev):
        return C[vi_prev - 1, nC[vi_prev - 1] - 1]

    def mark(vi, val):
        T[iTri[vi - 1, 0] - 1] = val
        T[iTri[vi - 1, niTri[vi - 1] - 1] - 1] = val

    vi_prev = 1
    vi = step(vi_prev)
    while abs(V[vi - 1, 1] - mesh.ymax) > tol:
        vi = step(vi_prev); mark(vi, 7); vi_prev = vi
    vi_NW = vi
    while abs(V[vi - 1, 0] - mesh.xmax) > tol:
        vi = step(vi_prev); mark(vi, 1); vi_prev = vi
    vi_NE = vi
    while abs(V[vi - 1, 1] - mesh.ymin) > tol:
        vi = step(vi_prev); mark(vi, 3); vi_prev = vi
    vi_SE = vi
    while abs(V[vi - 1, 0] - mesh.xmin) > tol:
        vi = step(vi_prev); mark(vi, 5); vi_prev = vi
    vi_SW = vi
    for v, val in ((vi_SW, 7), (vi_SE, 5), (vi_NE, 3), (vi_NW, 1)):
        T[iTri[v - 1, niTri[v - 1] - 1] - 1] = val
    return T


def _finish_geometry(mesh: Mesh):
    """Circumcentres, connection widths and Voronoi areas."""
    P = mesh.V
    tris = mesh.Tri - 1
    mesh.Tricc = np.asfortranarray(_circumcentres(P, tris))
    if mesh.nV > 5:
        mesh.Tri_edge_index = _tri_edge_index(mesh)
    nV, nC_mem = mesh.nV, mesh.nC_mem
    Cw = np.zeros((nV, nC_mem), np.float64, order="F")
    cc = mesh.Tricc
    for ci in range(int(mesh.nC.max())):
        has = mesh.nC > ci
        vi = np.nonzero(has)[0]
        vj = mesh.C[vi, ci] - 1
        # the (up to) two triangles containing edge (vi, vj): iTri(ci-1), iTri(ci)
        t_after = np.where(ci < mesh.niTri[vi], mesh.iTri[vi, np.minimum(ci, nC_mem - 1)], 0)
        prev = np.where(ci > 0, ci - 1, mesh.niTri[vi] - 1)
        edge = mesh.edge_index[vi] > 0
        t_before = np.where(edge & (ci == 0), 0, mesh.iTri[vi, prev])
        t1 = np.where(t_before > 0, t_before, t_after)
        t2 = np.where(t_before > 0, t_after, 0)
        two = t2 > 0
        w = np.zeros(vi.size)
        d = cc[t1[two] - 1] - cc[t2[two] - 1]
        w[two] = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
        one = ~two
        tt = t1[one] - 1
        te = mesh.Tri_edge_index[tt]
        c1 = cc[tt]
        ww = np.select([te == 1, te == 3, te == 5, te == 7],
                       [mesh.ymax - c1[:, 1], mesh.xmax - c1[:, 0],
                        c1[:, 1] - mesh.ymin, c1[:, 0] - mesh.xmin], default=np.nan)
        w[one] = np.maximum(0.0, ww)
        Cw[vi, ci] = w
        del vj
    mesh.Cw = Cw
    # Voronoi area as the sum of the kites spanned by vi and each cell edge
    A = np.zeros(nV)
    for ci in range(int(mesh.nC.max())):
        vi = np.nonzero(mesh.nC > ci)[0]
        vj = mesh.C[vi, ci] - 1
        dist = np.hypot(P[vj, 0] - P[vi, 0], P[vj, 1] - P[vi, 1])
        A[vi] += dist * Cw[vi, ci] / 4.0
    mesh.A = A


# boundary sides: (edge_index of the side, the corners on it, axis along it, fixed coordinate)
_SIDES = {1: (2, 8), 3: (2, 4), 5: (4, 6), 7: (6, 8)}


def _uncroppable_triangles(P, edge_index, simplices, xmin, xmax, ymin, ymax, tol):
    """Triangles whose circumcentre lies outside the domain beyond a side on
    which the triangle has no edge.  The reference crops a circumcentre only
    back across the side its triangle borders (crop_circumcenter,
    src/mesh_help_functions_module.f90:452-487) and stops on any other Voronoi
    vertex outside the domain (:240-247); its refinement never leaves such a
    triangle."""
    a, b, c = (P[simplices[:, k]] for k in range(3))
    d = 2.0 * (a[:, 0] * (b[:, 1] - c[:, 1]) + b[:, 0] * (c[:, 1] - a[:, 1]) + c[:, 0] * (a[:, 1] - b[:, 1]))
    sa, sb, sc = (np.einsum("ij,ij->i", q, q) for q in (a, b, c))
    ux = (sa * (b[:, 1] - c[:, 1]) + sb * (c[:, 1] - a[:, 1]) + sc * (a[:, 1] - b[:, 1])) / d
    uy = (sa * (c[:, 0] - b[:, 0]) + sb * (a[:, 0] - c[:, 0]) + sc * (b[:, 0] - a[:, 0])) / d
    ev = edge_index[simplices]
    bad = np.zeros(len(simplices), bool)
    for side, outside in ((1, uy > ymax + tol), (3, ux > xmax + tol), (5, uy < ymin - tol), (7, ux < xmin - tol)):
        on = np.isin(ev, (side,) + _SIDES[side]).sum(1) >= 2
        bad |= outside & ~on
    return np.nonzero(bad)[0]


def _split_encroached_boundary(P, edge_index, tri, xmin, xmax, ymin, ymax, max_rounds=20):
    """Ruppert's boundary-segment splitting, applied where the mesh needs it:
    a boundary segment whose diametral circle holds a vertex of an
    uncroppable triangle is split at its midpoint, and the points are
    triangulated again.  Once no boundary segment is encroached every
    circumcentre lies inside the domain, so the loop ends; meshes without such
    triangles are returned unchanged."""
    from scipy.spatial import Delaunay, cKDTree
    tol = ((xmax - xmin) + (ymax - ymin)) * 1e-9 / 2.0
    for _ in range(max_rounds):
        bad = _uncroppable_triangles(P, edge_index, tri.simplices, xmin, xmax, ymin, ymax, tol)
        if bad.size == 0:
            return P, edge_index, tri
        suspects = set(np.unique(tri.simplices[bad]).tolist())
        tree = cKDTree(P)
        new_p, new_e = [], []
        for side, corners in _SIDES.items():
            idx = np.nonzero(np.isin(edge_index, (side,) + corners))[0]
            axis = 0 if side in (1, 5) else 1
            idx = idx[np.argsort(P[idx, axis], kind="stable")]
            for i, j in zip(idx[:-1], idx[1:]):
                m = 0.5 * (P[i] + P[j])
                r = 0.5 * np.hypot(*(P[j] - P[i]))
                inside = [k for k in tree.query_ball_point(m, r * (1.0 - 1e-12)) if k not in (i, j)]
                if any(k in suspects for k in inside):
                    new_p.append(m)
                    new_e.append(side)
        if not new_p:
            raise ValueError("mesh has a circumcentre outside the domain that no boundary split removes")
        P = np.concatenate([P, np.array(new_p)])
        edge_index = np.concatenate([edge_index, np.array(new_e, edge_index.dtype)])
        tri = Delaunay(P)
    raise ValueError("boundary splitting did not converge")


def generate_mesh(res, n_target, seed=0, xmin=-800e3, xmax=800e3, ymin=-800e3,
                  ymax=800e3, jitter=0.15, nC_mem=NC_MEM, max_iter=3, smooth=2,
                  split_encroached=True) -> Mesh:
    """Seeded mesh with about ``n_target`` vertices following ``res(x, y)``.

    ``res`` gives the relative target edge length [m]; it is rescaled
    uniformly until the vertex count is within a few percent of the target.
    """
    from scipy.spatial import Delaunay

    # first guess from the point density of a triangular lattice, 2/(sqrt(3) h^2)
    gx, gy = np.meshgrid(np.linspace(xmin, xmax, 129), np.linspace(ymin, ymax, 129))
    r = res(gx.ravel(), gy.ravel())
    dens = np.mean(2.0 / (math.sqrt(3.0) * r ** 2)) * (xmax - xmin) * (ymax - ymin)
    scale = math.sqrt(dens / n_target)
    for _ in range(max_iter):
        rng = np.random.default_rng(seed)
        X, Y, bx, by, be = _place_points(res, xmin, xmax, ymin, ymax, scale, rng, jitter)
        n = X.size + bx.size + 4
        if abs(n - n_target) <= 0.03 * n_target:
            break
        scale *= math.sqrt(n / n_target)
    corners = np.array([[xmin, ymin], [xmax, ymin], [xmax, ymax], [xmin, ymax]])
    rest = np.concatenate([np.stack([bx, by], 1), np.stack([X, Y], 1)])
    rest_edge = np.concatenate([be, np.zeros(X.size, np.int32)])
    span = max(xmax - xmin, ymax - ymin)
    gx = np.clip(((rest[:, 0] - xmin) / span * 65535), 0, 65535).astype(np.int64)
    gy = np.clip(((rest[:, 1] - ymin) / span * 65535), 0, 65535).astype(np.int64)
    o = np.argsort(_hilbert_index(gx, gy), kind="stable")
    P = np.concatenate([corners, rest[o]])
    edge_index = np.concatenate([np.array([6, 4, 2, 8], np.int32), rest_edge[o]])
    tri = Delaunay(P)
    # Laplacian smoothing of the interior vertices (the reference smooths its
    # meshes with Lloyd iterations, src/mesh_creation_module.f90:743-792):
    # skewed triangles give second-derivative stencils whose centre weight has
    # the wrong sign, and over-relaxed SOR diverges on those rows.
    interior = edge_index == 0
    for _ in range(smooth):
        ptr, nbr = tri.vertex_neighbor_vertices
        cnt = np.diff(ptr)
        mean = np.stack([np.add.reduceat(P[nbr, 0], ptr[:-1]), np.add.reduceat(P[nbr, 1], ptr[:-1])], 1)
        mean /= cnt[:, None]
        P = np.where(interior[:, None], P + 0.6 * (mean - P), P)
        tri = Delaunay(P)
    if split_encroached:
        P, edge_index, tri = _split_encroached_boundary(P, edge_index, tri, xmin, xmax, ymin, ymax)
    if len(tri.coplanar):
        raise ValueError("triangulation dropped points")
    tris, nC, C, niTri, iTri, TriC = _build_topology(
        P, edge_index, tri.simplices.astype(np.int64), xmin, xmax, ymin, ymax, nC_mem)
    mesh = Mesh(xmin, xmax, ymin, ymax, nC_mem,
                np.asfortranarray(P), nC, C, niTri, iTri,
                edge_index.astype(np.int32), np.asfortranarray((tris + 1).astype(np.int32)),
                np.asfortranarray(TriC), None, None, None, None)
    _finish_geometry(mesh)
    return mesh
