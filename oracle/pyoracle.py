"""ctypes binding of the CPU oracle (test infrastructure only).

Loaded by tests/, the smoke check and bench.py's CPU-baseline leg; the
product package never imports this module.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from ufemism_amd import fields as F
from ufemism_amd.config import DynParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libufe_oracle.so")

def _count(fn):
    lib = ct.CDLL(LIB_PATH)
    f = getattr(lib, fn)
    f.restype = ct.c_char_p
    n = 0
    while f(n) is not None:
        n += 1
    return n


if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
N_MESH_D = _count("om_mesh_field_name_d")
N_MESH_I = _count("om_mesh_field_name_i")
N_ICE_D = _count("om_ice_field_name_d")
N_ICE_I = _count("om_ice_field_name_i")


class _Dims(ct.Structure):
    _fields_ = [("nV", ct.c_int), ("nTri", ct.c_int), ("nC_mem", ct.c_int), ("nAc", ct.c_int),
                ("nVAaAc", ct.c_int), ("xmin", ct.c_double), ("xmax", ct.c_double),
                ("ymin", ct.c_double), ("ymax", ct.c_double), ("colour_nV", ct.c_int * 5)]


class _Mesh(ct.Structure):
    _fields_ = [("d", _Dims), ("f", ct.c_void_p * N_MESH_D), ("i", ct.c_void_p * N_MESH_I)]


class _Ice(ct.Structure):
    _fields_ = [("f", ct.c_void_p * N_ICE_D), ("i", ct.c_void_p * N_ICE_I)]


class _Params(ct.Structure):
    _fields_ = [("nZ", ct.c_int), ("zeta", ct.c_double * 32), ("m_enh_sia", ct.c_double),
                ("m_enh_ssa", ct.c_double), ("SSA_RN_tol", ct.c_double),
                ("SSA_max_outer_loops", ct.c_int), ("SSA_max_residual_UV", ct.c_double),
                ("SSA_SOR_omega", ct.c_double), ("SSA_max_inner_loops", ct.c_int),
                ("use_analytical_GL_flux", ct.c_int), ("ssa_forced_zero", ct.c_int),
                ("zero_Hi_at_edges", ct.c_int), ("use_tau_c_override", ct.c_int)]


class SSAStats(ct.Structure):
    _fields_ = [("picard_iterations", ct.c_int), ("picard_converged", ct.c_int),
                ("resets", ct.c_int), ("total_sweeps", ct.c_int), ("last_RN", ct.c_double),
                ("last_max_residual", ct.c_double), ("sweeps", ct.c_int * 64),
                ("status", ct.c_int)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        _lib = ct.CDLL(LIB_PATH)
        for fn in ("om_mesh_field_name_d", "om_mesh_field_name_i", "om_ice_field_name_d",
                   "om_ice_field_name_i"):
            getattr(_lib, fn).restype = ct.c_char_p
        _lib.om_vertical_average.restype = ct.c_double
    return _lib


def _names(fn, n):
    lib = load()
    out = []
    for k in range(n):
        s = getattr(lib, fn)(k)
        out.append(s.decode())
    assert getattr(lib, fn)(n) is None
    return out


def params_struct(p: DynParams) -> _Params:
    s = _Params()
    s.nZ = p.nZ
    for k, z in enumerate(p.zeta):
        s.zeta[k] = float(z)
    s.m_enh_sia, s.m_enh_ssa = p.m_enh_sia, p.m_enh_ssa
    s.SSA_RN_tol, s.SSA_max_outer_loops = p.SSA_RN_tol, p.SSA_max_outer_loops
    s.SSA_max_residual_UV, s.SSA_SOR_omega = p.SSA_max_residual_UV, p.SSA_SOR_omega
    s.SSA_max_inner_loops = p.SSA_max_inner_loops
    s.use_analytical_GL_flux = int(p.use_analytical_GL_flux)
    s.ssa_forced_zero = int(p.ssa_forced_zero)
    s.zero_Hi_at_edges = int(p.zero_Hi_at_edges)
    s.use_tau_c_override = int(p.use_tau_c_override)
    return s


class OracleMesh:
    """A primary mesh plus the secondary arrays the oracle builds from it."""

    def __init__(self, mesh, nZ=15, geometry=True):
        self.lib = load()
        self.mesh = mesh
        self.nZ = nZ
        nV, nC_mem = mesh.nV, mesh.nC_mem
        self.nAc = int(mesh.nC[:].sum() // 2)
        self.nVAaAc = nV + self.nAc
        self.arr = {
            "V": mesh.V, "A": mesh.A, "Cw": mesh.Cw, "nC": mesh.nC, "C": mesh.C,
            "niTri": mesh.niTri, "iTri": mesh.iTri, "edge_index": mesh.edge_index,
            "Tri": mesh.Tri,
        }
        if geometry:
            # Tricc, A and Cw from the primary arrays, as the reference computes
            # them after creating or reading a mesh
            self.geometry = mesh_geometry(mesh)
            for k in ("Tri", "Tricc", "A", "Cw", "TriA", "R"):
                self.arr[k] = self.geometry[k]
        self.arr.update(F.alloc(F.MESH_SECONDARY, nV, self.nAc, nC_mem, nZ))
        s = _Mesh()
        s.d.nV, s.d.nTri, s.d.nC_mem = nV, mesh.nTri, nC_mem
        s.d.nAc, s.d.nVAaAc = self.nAc, self.nVAaAc
        s.d.xmin, s.d.xmax, s.d.ymin, s.d.ymax = mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax
        for k, name in enumerate(_names("om_mesh_field_name_d", N_MESH_D)):
            a = self.arr[name] = np.asfortranarray(self.arr[name], dtype=np.float64)
            assert a.dtype == np.float64 and a.flags.f_contiguous, name
            s.f[k] = a.ctypes.data
        for k, name in enumerate(_names("om_mesh_field_name_i", N_MESH_I)):
            a = self.arr[name] = np.asfortranarray(self.arr[name], dtype=np.int32)
            assert a.dtype == np.int32 and a.flags.f_contiguous, name
            s.i[k] = a.ctypes.data
        self.s = s
        self.colour_status = None

    @property
    def colour_nV(self):
        return list(self.s.d.colour_nV)

    def build(self, colour=True):
        r = self.lib.om_make_Ac_mesh(ct.byref(self.s))
        if r != 0:
            raise RuntimeError(f"om_make_Ac_mesh failed ({r})")
        self.lib.om_get_neighbour_functions(ct.byref(self.s))
        if colour:
            self.colour_status = self.lib.om_calculate_five_colouring_AaAc(ct.byref(self.s))
        return self

    def __getitem__(self, k):
        return self.arr[k]


class OracleIce:
    """Host arrays of one ice state, bound to an OracleMesh."""

    def __init__(self, om: OracleMesh):
        self.om = om
        self.lib = om.lib
        self.arr = F.alloc(F.ICE_FIELDS, om.mesh.nV, om.nAc, om.mesh.nC_mem, om.nZ)
        s = _Ice()
        for k, name in enumerate(_names("om_ice_field_name_d", N_ICE_D)):
            s.f[k] = self.arr[name].ctypes.data
        for k, name in enumerate(_names("om_ice_field_name_i", N_ICE_I)):
            s.i[k] = self.arr[name].ctypes.data
        self.s = s

    def __getitem__(self, k):
        return self.arr[k]

    def set(self, **kw):
        for k, v in kw.items():
            self.arr[k][...] = v

    def _m(self):
        return ct.byref(self.om.s)

    def update_general_ice_model_data(self):
        self.lib.om_update_general_ice_model_data(self._m(), ct.byref(self.s))

    def ice_physical_properties(self, p: DynParams, mode=1):
        ps = params_struct(p)
        self.lib.om_ice_physical_properties(self._m(), ct.byref(self.s), ct.byref(ps), ct.c_int(mode))

    def solve_SIA(self, p: DynParams):
        ps = params_struct(p)
        self.lib.om_solve_SIA(self._m(), ct.byref(self.s), ct.byref(ps))

    def solve_SSA(self, p: DynParams) -> SSAStats:
        ps = params_struct(p)
        st = SSAStats()
        self.lib.om_solve_SSA(self._m(), ct.byref(self.s), ct.byref(ps), ct.byref(st))
        return st

    def solve_SIA_3D(self, p: DynParams):
        ps = params_struct(p)
        self.lib.om_solve_SIA_3D(self._m(), ct.byref(self.s), ct.byref(ps))

    def critical_timesteps(self, p: DynParams):
        ps = params_struct(p)
        out = (ct.c_double * 3)()
        self.lib.om_critical_timesteps(self._m(), ct.byref(self.s), ct.byref(ps), out)
        return tuple(out)

    def calculate_ice_thickness_change(self, SMB, BMB, dt, mask_noice, p: DynParams):
        ps = params_struct(p)
        SMB = np.ascontiguousarray(SMB, dtype=np.float64)
        BMB = np.ascontiguousarray(BMB, dtype=np.float64)
        mask = np.ascontiguousarray(mask_noice, dtype=np.int32)
        self.lib.om_calculate_ice_thickness_change(
            self._m(), ct.byref(self.s), SMB.ctypes.data_as(ct.c_void_p),
            BMB.ctypes.data_as(ct.c_void_p), ct.c_double(dt),
            mask.ctypes.data_as(ct.c_void_p), ct.byref(ps))


def op(name, om: OracleMesh, *arrays, nz=None):
    """Call one operator routine on contiguous float64 arrays (in place)."""
    lib = load()
    args = [ct.byref(om.s)] + [a.ctypes.data_as(ct.c_void_p) for a in arrays]
    if nz is not None:
        args.append(ct.c_int(nz))
    getattr(lib, "om_" + name)(*args)


def set_threads(n):
    return load().om_set_threads(int(n))


def run_step(wl, threads=None):
    """The same single dynamics step as ufemism_amd.dynamics.Step, on the oracle."""
    from ufemism_amd.dynamics import OUTPUTS, initial_fields
    if threads:
        set_threads(threads)
    om = OracleMesh(wl.mesh).build()
    if om.colour_status != 0:
        raise RuntimeError(f"oracle colouring status {om.colour_status}")
    ice = OracleIce(om)
    for k, v in initial_fields(wl, om["VAaAc"], wl.params.nZ).items():
        if k in ice.arr:
            ice[k][...] = v
    ice.update_general_ice_model_data()
    if wl.params.physical_properties:
        ice.ice_physical_properties(wl.params, wl.params.physical_properties)
    ice.solve_SIA(wl.params)
    st = ice.solve_SSA(wl.params)
    ice.calculate_ice_thickness_change(wl.SMB, wl.BMB, wl.dt,
                                       np.zeros(wl.mesh.nV, np.int32), wl.params)
    out = {k: ice[k].copy() for k in OUTPUTS}
    out["sweeps"] = list(st.sweeps)[: st.picard_iterations]
    out["picard"] = st.picard_iterations
    out["stats"] = st
    out["mesh"] = om
    out["ice"] = ice
    return out


def mesh_from_arrays(mesh, arrays, colour_nV):
    """OracleMesh filled with secondary arrays built elsewhere (used only to
    time the oracle's sweeps on a mesh too large for its literal colouring)."""
    om = OracleMesh(mesh, geometry=False)
    for k, a in om.arr.items():
        if k in arrays and k not in ("V", "A", "Cw", "nC", "C", "niTri", "iTri", "edge_index", "Tri"):
            a[...] = arrays[k]
    for k in range(5):
        om.s.d.colour_nV[k] = int(colour_nV[k])
    return om


def time_sor_sweeps(om, fields, p: DynParams, sweeps, threads=None):
    """Run `sweeps` oracle SOR sweeps of solve_SSA_linearised on the given
    AaAc fields (original order) and return the wall time in seconds."""
    import time
    if threads:
        set_threads(threads)
    ice = OracleIce(om)
    for k, v in fields.items():
        ice[k][...] = v
    q = DynParams(**{f.name: getattr(p, f.name) for f in __import__("dataclasses").fields(p)})
    q.SSA_max_inner_loops = int(sweeps)
    q.SSA_max_residual_UV = 0.0
    ps = params_struct(q)
    did = ct.c_int(0)
    lm = ct.c_double(0)
    ns = ct.c_int(0)
    t0 = time.perf_counter()
    load().om_solve_SSA_linearised(ct.byref(om.s), ct.byref(ice.s), ct.byref(ps), ct.byref(did),
                                   ct.byref(lm), ct.byref(ns))
    dt = time.perf_counter() - t0
    return dt, ns.value


def colour_graph(adj, nC_mem=16):
    """Run the oracle's five-colouring on an arbitrary graph given as 1-based
    adjacency lists (one list per point); returns (status, colour array)."""
    n = len(adj)
    s = _Mesh()
    s.d.nV, s.d.nAc, s.d.nVAaAc, s.d.nC_mem = n, 0, n, nC_mem
    nC = np.array([len(a) for a in adj], np.int32)
    C = np.zeros((n, nC_mem), np.int32, order="F")
    for i, a in enumerate(adj):
        C[i, :len(a)] = a
    colour = np.zeros(n, np.int32)
    colour_vi = np.zeros((n, 5), np.int32, order="F")
    keep = {"nCAaAc": nC, "CAaAc": C, "colour": colour, "colour_vi": colour_vi}
    for k, name in enumerate(_names("om_mesh_field_name_i", N_MESH_I)):
        if name in keep:
            s.i[k] = keep[name].ctypes.data
    st = load().om_calculate_five_colouring_AaAc(ct.byref(s))
    return st, colour


class _Geom(ct.Structure):
    _fields_ = [("nV", ct.c_int), ("nTri", ct.c_int), ("nC_mem", ct.c_int),
                ("xmin", ct.c_double), ("xmax", ct.c_double), ("ymin", ct.c_double),
                ("ymax", ct.c_double), ("tol_dist", ct.c_double),
                ("V", ct.c_void_p), ("nC", ct.c_void_p), ("C", ct.c_void_p), ("niTri", ct.c_void_p),
                ("iTri", ct.c_void_p), ("edge_index", ct.c_void_p), ("Tri", ct.c_void_p),
                ("Tri_edge_index", ct.c_void_p), ("Tricc", ct.c_void_p)]


def _geom_struct(mesh, Tri, Tricc):
    g = _Geom()
    g.nV, g.nTri, g.nC_mem = mesh.nV, mesh.nTri, mesh.nC_mem
    g.xmin, g.xmax, g.ymin, g.ymax = mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax
    g.tol_dist = mesh.tol_dist
    keep = []
    for name, a in (("V", mesh.V), ("nC", mesh.nC), ("C", mesh.C), ("niTri", mesh.niTri),
                    ("iTri", mesh.iTri), ("edge_index", mesh.edge_index),
                    ("Tri_edge_index", mesh.Tri_edge_index)):
        a = np.asfortranarray(a, dtype=np.float64 if name == "V" else np.int32)
        keep.append(a)
        setattr(g, name, a.ctypes.data)
    g.Tri, g.Tricc = Tri.ctypes.data, Tricc.ctypes.data
    return g, keep


def mesh_geometry(mesh, compute_Tricc=True):
    """The reference's primary-mesh geometry on a mesh (src/mesh_help_functions_module.f90):
    Tricc (update_triangle_circumcenter, or the mesh's own when compute_Tricc is False),
    Voronoi areas A, connection widths Cw, triangle areas TriA and resolution R."""
    lib = load()
    Tri = np.array(mesh.Tri, dtype=np.int32, order="F")
    Tricc = (np.zeros((mesh.nTri, 2), order="F") if compute_Tricc
             else np.array(mesh.Tricc, dtype=np.float64, order="F"))
    g, keep = _geom_struct(mesh, Tri, Tricc)
    out = {"Tri": Tri, "Tricc": Tricc}
    rot = ct.c_int(0)
    if compute_Tricc:
        out["invalid_cc"] = lib.om_update_triangle_circumcenters(ct.byref(g), ct.byref(rot))
    out["rotated"] = rot.value
    A = np.zeros(mesh.nV)
    Cw = np.zeros((mesh.nV, mesh.nC_mem), order="F")
    TriA = np.zeros(mesh.nTri)
    R = np.zeros(mesh.nV)
    rmin, rmax = ct.c_double(), ct.c_double()
    out["outside"] = lib.om_find_Voronoi_cell_areas(ct.byref(g), A.ctypes.data_as(ct.c_void_p))
    out["cw_fails"] = lib.om_find_connection_widths(ct.byref(g), Cw.ctypes.data_as(ct.c_void_p))
    lib.om_find_triangle_areas(ct.byref(g), TriA.ctypes.data_as(ct.c_void_p))
    lib.om_determine_mesh_resolution(ct.byref(g), R.ctypes.data_as(ct.c_void_p), ct.byref(rmin),
                                     ct.byref(rmax))
    out.update(A=A, Cw=Cw, TriA=TriA, R=R, resolution_min=rmin.value, resolution_max=rmax.value)
    del keep
    return out


def voronoi_cell_vertices(mesh, Tricc, vi):
    """Clamped Voronoi cell vertices of vertex vi (1-based), as the reference lists them."""
    lib = load()
    Tri = np.array(mesh.Tri, dtype=np.int32, order="F")
    Tricc = np.asfortranarray(Tricc, dtype=np.float64)
    g, keep = _geom_struct(mesh, Tri, Tricc)
    mx = 2 * mesh.nC_mem + 4
    Vor = np.zeros(2 * mx)
    out = ct.c_int(0)
    n = lib.om_find_Voronoi_cell_vertices(ct.byref(g), int(vi), Vor.ctypes.data_as(ct.c_void_p), mx,
                                          ct.byref(out))
    del keep
    return Vor[: 2 * min(n, mx)].reshape(-1, 2), out.value


def upwind_derivatives_3D(om: OracleMesh, R, U, V, d):
    """get_upwind_derivative_vertex_3D on every vertex and layer of d (nV, nz),
    with the triangle neighbour functions of the mesh; returns (ddx, ddy, missing)."""
    lib = load()
    nT = om.mesh.nTri
    NxT = np.zeros((nT, 3), order="F")
    NyT = np.zeros((nT, 3), order="F")
    lib.om_triangle_neighbour_functions(ct.byref(om.s), NxT.ctypes.data_as(ct.c_void_p),
                                        NyT.ctypes.data_as(ct.c_void_p))
    U, V, d, R = (np.asfortranarray(a, dtype=np.float64) for a in (U, V, d, R))
    nz = d.shape[1]
    ddx = np.zeros_like(d, order="F")
    ddy = np.zeros_like(d, order="F")
    missing = lib.om_upwind_derivatives_3D(ct.byref(om.s), *(a.ctypes.data_as(ct.c_void_p) for a in (R, NxT, NyT, U, V, d)),
                                           ct.c_int(nz), ddx.ctypes.data_as(ct.c_void_p),
                                           ddy.ctypes.data_as(ct.c_void_p))
    return ddx, ddy, missing, NxT, NyT


def run_dynamics(wl, t_end, clock, dt_first=0.0):
    """The same time loop as ufemism_amd.dynamics.run_dynamics, on the oracle;
    returns (log, final fields)."""
    from ufemism_amd.dynamics import initial_fields
    om = OracleMesh(wl.mesh).build()
    ice = OracleIce(om)
    for k, v in initial_fields(wl, om["VAaAc"], wl.params.nZ).items():
        if k in ice.arr:
            ice[k][...] = v
    p = wl.params
    noice = np.zeros(wl.mesh.nV, np.int32)
    log = []
    clock.dt = dt_first
    while clock.time < t_end:
        ice.calculate_ice_thickness_change(wl.SMB, wl.BMB, clock.dt, noice, p)
        ice.update_general_ice_model_data()
        if p.physical_properties:
            ice.ice_physical_properties(p, p.physical_properties)
        st = None
        if clock.do["SIA"]:
            ice.solve_SIA(p)
            clock.t0["SIA"] = clock.time
        if clock.do["SSA"]:
            st = ice.solve_SSA(p)
            clock.t0["SSA"] = clock.time
        for a in ("climate", "SMB", "BMB"):
            if clock.do[a]:
                clock.t0[a] = clock.time
        if clock.do["thermo"]:
            ice.solve_SIA_3D(p)
            clock.t0["thermo"] = clock.time
        if clock.do["output"]:
            clock.t0["output"] = clock.time
        if clock.do["ELRA"]:
            clock.t0["ELRA"] = clock.time
        clock.determine_timesteps_and_actions(ice.critical_timesteps(p), t_end)
        log.append((clock.time, clock.dt, dict(clock.do), st))
    return log, ice, om


# ---- remapping (ufe_oracle_remap.c), src/mesh_mapping_module.f90 ----------

def _geom_of(mesh, geo):
    """om_geom of a mesh with its reference geometry (Tricc, Tri from mesh_geometry)."""
    Tri = np.array(geo["Tri"], dtype=np.int32, order="F")
    Tricc = np.array(geo["Tricc"], dtype=np.float64, order="F")
    g, keep = _geom_struct(mesh, Tri, Tricc)
    keep += [g, Tri, Tricc]
    return g, keep


def geometric_centres(mesh, geo):
    """find_Voronoi_cell_geometric_centres: VorGC (nV, 2)."""
    lib = load()
    g, keep = _geom_of(mesh, geo)
    A = np.ascontiguousarray(geo["A"])
    GC = np.zeros((mesh.nV, 2), order="F")
    lib.om_voronoi_geometric_centres(ct.byref(g), A.ctypes.data_as(ct.c_void_p), GC.ctypes.data_as(ct.c_void_p))
    del keep
    return GC


def cell_polygon(mesh, geo, vi):
    lib = load()
    g, keep = _geom_of(mesh, geo)
    P = np.zeros(2 * 140)
    n = lib.om_cell_polygon(ct.byref(g), int(vi), P.ctypes.data_as(ct.c_void_p), 136)
    del keep
    return P[: 2 * n].reshape(-1, 2)


class ConservativeMap:
    """Entries start[vd]:end[vd] (0-based, end exclusive) of each destination
    vertex: source vertex vi (1-based) and weights w0, w1x, w1y."""

    def __init__(self, ptr, vi, w0, w1x, w1y):
        self.start = ptr[:-1].copy()
        self.end = ptr[1:].copy()
        self.vi, self.w0, self.w1x, self.w1y = vi, w0, w1x, w1y


def remap_conservative_weights(src, src_geo, dst, dst_geo):
    lib = load()
    gs, ks = _geom_of(src, src_geo)
    gd, kd = _geom_of(dst, dst_geo)
    As = np.ascontiguousarray(src_geo["A"])
    Ad = np.ascontiguousarray(dst_geo["A"])
    GC = np.asfortranarray(geometric_centres(src, src_geo))
    ptr = np.zeros(dst.nV + 1, np.int32)
    P = lambda a: a.ctypes.data_as(ct.c_void_p)
    n = lib.om_remap_conservative(ct.byref(gs), P(As), P(GC), ct.byref(gd), P(Ad), P(ptr), None, None, None, None, 0)
    if n < 0:
        raise RuntimeError("om_remap_conservative: polygon overflow")
    vi = np.zeros(n, np.int32)
    w0, w1x, w1y = np.zeros(n), np.zeros(n), np.zeros(n)
    lib.om_remap_conservative(ct.byref(gs), P(As), P(GC), ct.byref(gd), P(Ad), P(ptr), P(vi), P(w0), P(w1x),
                              P(w1y), ct.c_int(n))
    del ks, kd
    return ConservativeMap(ptr, vi, w0, w1x, w1y)


def remap_trilin_nearest(src, src_geo, dst):
    """create_remapping_arrays_trilin / _nearest_neighbour: (tri_vi (nV,3) 1-based,
    tri_w (nV,3), nn_vi (nV) 1-based, missing)."""
    lib = load()
    gs, ks = _geom_of(src, src_geo)
    gd, kd = _geom_of(dst, {"Tri": dst.Tri, "Tricc": np.zeros((dst.nTri, 2))})
    TriC = np.asfortranarray(src.TriC, dtype=np.int32)
    tvi = np.zeros((dst.nV, 3), np.int32, order="F")
    tw = np.zeros((dst.nV, 3), order="F")
    nn = np.zeros(dst.nV, np.int32)
    P = lambda a: a.ctypes.data_as(ct.c_void_p)
    miss = lib.om_remap_trilin_nearest(ct.byref(gs), P(TriC), ct.byref(gd), P(tvi), P(tw), P(nn))
    del ks, kd
    return tvi, tw, nn, miss


def apply_conservative(cmap, d_src, ddx=None, ddy=None):
    """remap_cons_1st_order_2D (src/mesh_mapping_module.f90:3964-3986) or, with
    the source derivatives, remap_cons_2nd_order_2D (:4013-4047), summed in
    entry order."""
    nd = len(cmap.start)
    out = np.zeros(nd)
    for vd in range(nd):
        s = 0.0
        for e in range(cmap.start[vd], cmap.end[vd]):
            v = cmap.vi[e] - 1
            if ddx is None:
                s = s + (d_src[v] * cmap.w0[e])
            else:
                s = s + (d_src[v] * cmap.w0[e]) + (ddx[v] * cmap.w1x[e]) + (ddy[v] * cmap.w1y[e])
        out[vd] = s
    return out


def _tri_area(a, b, c):
    return abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])) / 2.0


def check_conservative(cmap, src, src_geo, src_om, dst):
    """check_if_remapping_is_conservative (src/mesh_mapping_module.f90:2416-2566):
    remap ones with the 2nd-order weights; where an interior destination vertex
    is off by more than 0.1, replace its entries by triangle weights of the
    containing source triangle (keeping the reference's weights for two
    entries, which repeat Aa/(Aa+Ab) when C is the smallest).  Raises where the
    reference stops (more than ceil(nV/100) such vertices).  Returns the number
    of replaced vertices."""
    lib = load()
    ones = np.ones(src.nV)
    ddx, ddy = np.zeros(src.nV), np.zeros(src.nV)
    op("get_mesh_derivatives", src_om, ones, ddx, ddy)
    d = apply_conservative(cmap, ones, ddx, ddy)
    d[dst.edge_index > 0] = 1.0
    gs, ks = _geom_of(src, src_geo)
    TriC = np.asfortranarray(src.TriC, dtype=np.int32)
    n_wrong = 0
    for vd in np.nonzero(np.abs(1.0 - d) > 1e-1)[0]:
        n_wrong += 1
        p = (ct.c_double * 2)(dst.V[vd, 0], dst.V[vd, 1])
        t = lib.om_find_containing_triangle(ct.byref(gs), TriC.ctypes.data_as(ct.c_void_p), p, 1)
        tri = src_geo["Tri"][t - 1]
        pa, pb, pc = (src.V[tri[k] - 1] for k in range(3))
        pp = dst.V[vd]
        At, Aa, Ab, Ac = _tri_area(pa, pb, pc), _tri_area(pb, pc, pp), _tri_area(pc, pa, pp), _tri_area(pa, pb, pp)
        s, e = cmap.start[vd], cmap.end[vd]
        cmap.w1x[s:e] = cmap.w1y[s:e] = 0.0
        if e - s == 2:
            if Aa <= Ab and Aa <= Ac:
                vals = ([tri[1], tri[2]], [Ab / (Ab + Ac), Ac / (Ab + Ac)])
            elif Ab <= Aa and Ab <= Ac:
                vals = ([tri[0], tri[2]], [Aa / (Aa + Ac), Ac / (Aa + Ac)])
            else:
                vals = ([tri[0], tri[1]], [Aa / (Aa + Ab), Aa / (Aa + Ab)])
        elif e - s == 1:
            k = 0 if (Aa >= Ab and Aa >= Ac) else (1 if (Ab >= Aa and Ab >= Ac) else 2)
            vals = ([tri[k]], [1.0])
        else:
            vals = ([tri[0], tri[1], tri[2]], [Aa / At, Ab / At, Ac / At])
        n = len(vals[0])
        cmap.vi[s:s + n] = vals[0]
        cmap.w0[s:s + n] = vals[1]
        cmap.end[vd] = s + n
    del ks
    if n_wrong > int(np.ceil(dst.nV / 100.0)):
        raise RuntimeError(f"2nd order conservative remapping goes wrong for {n_wrong} out of {dst.nV} vertices")
    return n_wrong
