"""ctypes binding of libufe_dyn.so (the C-ABI declared in include/ufe_dyn.h).

The library is the product: every call here goes to the HIP implementation.
If the shared object is missing, importing works but ``load()`` raises, so
no caller can silently fall back to a CPU path.
"""
from __future__ import annotations

import ctypes as ct
import os
import warnings

import numpy as np

from . import fields as F
from .config import DynParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libufe_dyn.so")

UFE_OK, UFE_SOR_NOT_CONVERGED, UFE_SSA_RESET, UFE_SSA_UNSTABLE = 0, 1, 2, 3
UFE_COLOURING_IDENTIFY, UFE_COLOURING_FAILED, UFE_BAD_INPUT = 4, 5, 6
UFE_HIP_ERROR, UFE_NOT_SUPPORTED, UFE_HOST_ERROR = 7, 8, 9
UFE_FP64, UFE_MIXED, UFE_FP64_KRYLOV = 0, 1, 2
UFE_PRECOND_SOR, UFE_PRECOND_MULTIGRID = 0, 1

STATUS_NAMES = {0: "OK", 1: "SOR_NOT_CONVERGED", 2: "SSA_RESET", 3: "SSA_UNSTABLE",
                4: "COLOURING_IDENTIFY", 5: "COLOURING_FAILED", 6: "BAD_INPUT",
                7: "HIP_ERROR", 8: "NOT_SUPPORTED", 9: "HOST_ERROR"}

# every symbol include/ufe_dyn.h declares
EXPORTS = [
    "ufe_ctx_create", "ufe_ctx_destroy", "ufe_set_params", "ufe_last_error", "ufe_version",
    "ufe_mesh_build", "ufe_mesh_upload", "ufe_mesh_sizes", "ufe_mesh_geometry",
    "ufe_remap_build", "ufe_remap_field", "ufe_remap_ice_model", "ufe_remap_export",
    "ufe_multigrid_info_get",
    "ufe_field_count", "ufe_field_name", "ufe_field_id", "ufe_field_info", "ufe_field_set",
    "ufe_field_get", "ufe_field_set_device", "ufe_field_get_device",
    "ufe_update_general_ice_model_data", "ufe_ice_physical_properties", "ufe_solve_SIA", "ufe_solve_SSA",
    "ufe_calculate_ice_thickness_change", "ufe_solve_SIA_3D", "ufe_critical_timesteps",
    "ufe_op_upwind_derivatives_3D", "ufe_ssa_prepare", "ufe_ssa_linear_solve",
    "ufe_ssa_residual", "ufe_ssa_sor_solve_noalloc",
    "ufe_op_derivatives_Aa", "ufe_op_curvatures_Aa", "ufe_op_derivatives_Aa_3D",
    "ufe_op_neumann_Aa", "ufe_op_neumann_Aa_3D", "ufe_op_derivatives_Ac",
    "ufe_op_derivatives_AaAc", "ufe_op_curvatures_AaAc", "ufe_op_neumann_AaAc",
    "ufe_op_map_Aa_to_Ac", "ufe_op_map_Aa_to_Ac_3D", "ufe_op_map_Ac_to_Aa",
    "ufe_op_map_Ac_to_Aa_3D", "ufe_op_rotate_xy_to_po",
    "ufe_device_alloc", "ufe_device_free", "ufe_memcpy_h2d", "ufe_memcpy_d2h", "ufe_synchronize",
    "ufe_field_fill", "ufe_profile_enable", "ufe_profile_read", "ufe_sweep_bytes",
    "ufe_ssa_adjoint_forward", "ufe_ssa_adjoint_backward",
    "ufe_batch_define", "ufe_batch_linear_solve", "ufe_batch_solve_SSA",
    "ufe_sweep_autotune", "ufe_sweep_tuning_set",
]

_P = ct.c_void_p


class MeshDesc(ct.Structure):
    _fields_ = [("nV", ct.c_int32), ("nTri", ct.c_int32), ("nC_mem", ct.c_int32), ("nAc", ct.c_int32),
                ("xmin", ct.c_double), ("xmax", ct.c_double), ("ymin", ct.c_double), ("ymax", ct.c_double),
                ("V", _P), ("nC", _P), ("C", _P), ("niTri", _P), ("iTri", _P), ("edge_index", _P),
                ("Tri", _P), ("A", _P), ("Cw", _P),
                ("Nx", _P), ("Ny", _P), ("Nxx", _P), ("Nxy", _P), ("Nyy", _P),
                ("iAci", _P), ("Aci", _P), ("VAc", _P), ("Nx_Ac", _P), ("Ny_Ac", _P), ("No_Ac", _P),
                ("Np_Ac", _P), ("edge_index_Ac", _P), ("VAaAc", _P), ("nCAaAc", _P), ("CAaAc", _P),
                ("Nx_AaAc", _P), ("Ny_AaAc", _P), ("Nxx_AaAc", _P), ("Nxy_AaAc", _P), ("Nyy_AaAc", _P),
                ("colour", _P), ("colour_vi", _P), ("colour_nV", ct.c_int32 * 5)]


class MeshGeometryDesc(ct.Structure):
    _fields_ = [("nV", ct.c_int32), ("nTri", ct.c_int32), ("nC_mem", ct.c_int32),
                ("xmin", ct.c_double), ("xmax", ct.c_double), ("ymin", ct.c_double), ("ymax", ct.c_double),
                ("tol_dist", ct.c_double),
                ("V", _P), ("nC", _P), ("C", _P), ("niTri", _P), ("iTri", _P), ("edge_index", _P),
                ("Tri", _P), ("Tri_edge_index", _P), ("compute_Tricc", ct.c_int32),
                ("Tricc", _P), ("A", _P), ("Cw", _P), ("TriA", _P), ("R", _P)]


class MeshGeometryStats(ct.Structure):
    _fields_ = [("resolution_min", ct.c_double), ("resolution_max", ct.c_double),
                ("area_error", ct.c_double), ("rotated_triangles", ct.c_int32),
                ("invalid_circumcentres", ct.c_int32), ("voronoi_outside", ct.c_int32),
                ("connection_failures", ct.c_int32)]


class RemapStats(ct.Structure):
    _fields_ = [("n_entries", ct.c_int32), ("n_fallback", ct.c_int32), ("trilin_missing", ct.c_int32),
                ("seconds", ct.c_double)]


UFE_REMAP_TRILIN, UFE_REMAP_NEAREST_NEIGHBOUR, UFE_REMAP_CONS_1ST_ORDER, UFE_REMAP_CONS_2ND_ORDER = 0, 1, 2, 3
REMAP_METHODS = {"trilin": 0, "nearest_neighbour": 1, "cons_1st_order": 2, "cons_2nd_order": 3}


def geometry_desc(ma):
    """MeshGeometryDesc of a built mesh (MeshArrays with the geometry of
    Context.mesh_build); the returned tuple keeps its arrays alive."""
    m = ma.mesh
    g = MeshGeometryDesc()
    g.nV, g.nTri, g.nC_mem = m.nV, m.nTri, m.nC_mem
    g.xmin, g.xmax, g.ymin, g.ymax = m.xmin, m.xmax, m.ymin, m.ymax
    g.tol_dist = m.tol_dist
    keep = []
    for name, a, dt in (("V", m.V, np.float64), ("nC", m.nC, np.int32), ("C", m.C, np.int32),
                        ("niTri", m.niTri, np.int32), ("iTri", m.iTri, np.int32),
                        ("edge_index", m.edge_index, np.int32), ("Tri_edge_index", m.Tri_edge_index, np.int32),
                        ("Tri", ma["Tri"], np.int32), ("Tricc", ma["Tricc"], np.float64),
                        ("A", ma["A"], np.float64)):
        a = np.asfortranarray(a, dtype=dt)
        keep.append(a)
        setattr(g, name, a.ctypes.data)
    g.compute_Tricc = 0
    return g, keep


class Params(ct.Structure):
    _fields_ = [("nZ", ct.c_int32), ("zeta", ct.c_double * 32), ("m_enh_sia", ct.c_double),
                ("m_enh_ssa", ct.c_double), ("SSA_RN_tol", ct.c_double),
                ("SSA_max_outer_loops", ct.c_int32), ("SSA_max_residual_UV", ct.c_double),
                ("SSA_SOR_omega", ct.c_double), ("SSA_max_inner_loops", ct.c_int32),
                ("use_analytical_GL_flux", ct.c_int32), ("ssa_forced_zero", ct.c_int32),
                ("zero_Hi_at_edges", ct.c_int32), ("use_tau_c_override", ct.c_int32),
                ("precision", ct.c_int32), ("sweep_batch", ct.c_int32),
                ("SSA_linear_rel_tol", ct.c_double), ("SSA_gmres_restart", ct.c_int32),
                ("SSA_precond_sweeps", ct.c_int32), ("physical_properties", ct.c_int32),
                ("SSA_precond_type", ct.c_int32)]


class SSAStats(ct.Structure):
    _fields_ = [("picard_iterations", ct.c_int32), ("picard_converged", ct.c_int32),
                ("resets", ct.c_int32), ("total_sweeps", ct.c_int32), ("last_RN", ct.c_double),
                ("last_max_residual", ct.c_double), ("sweeps", ct.c_int32 * 64),
                ("status", ct.c_int32), ("sor_seconds", ct.c_double), ("total_seconds", ct.c_double)]


class LinSolveOpts(ct.Structure):
    _fields_ = [("rel_tol", ct.c_double), ("max_sweeps", ct.c_int32), ("precision", ct.c_int32),
                ("omega", ct.c_double), ("check_interval", ct.c_int32), ("use_graph", ct.c_int32),
                ("gmres_restart", ct.c_int32), ("precond_sweeps", ct.c_int32),
                ("orth_passes", ct.c_int32), ("method", ct.c_int32),
                ("precond_full_sweeps", ct.c_int32), ("precond_neumann_every_sweep", ct.c_int32),
                ("precond_type", ct.c_int32)]


class MultigridInfo(ct.Structure):
    _fields_ = [("levels", ct.c_int32), ("points", ct.c_int32 * 16), ("couplings", ct.c_int64 * 16),
                ("direct", ct.c_int32), ("kstep", ct.c_int32 * 16), ("setup_seconds", ct.c_double),
                ("tail_level", ct.c_int32), ("fused", ct.c_int32 * 16), ("fuse_list_max", ct.c_int32 * 16)]

    def as_dict(self):
        L = self.levels
        return {"levels": L, "points": list(self.points)[:L], "couplings": list(self.couplings)[:L],
                "direct": bool(self.direct), "kstep": [bool(k) for k in list(self.kstep)[:L]],
                "setup_seconds": self.setup_seconds, "tail_level": self.tail_level,
                "fused": list(self.fused)[:L], "fuse_list_max": list(self.fuse_list_max)[:L]}


class KernelProfile(ct.Structure):
    _fields_ = [("launches", ct.c_int64), ("seconds", ct.c_double),
                ("algorithmic_bytes", ct.c_double), ("rows", ct.c_int64), ("total_launches", ct.c_int64)]


class LinSolveStats(ct.Structure):
    _fields_ = [("sweeps", ct.c_int32), ("refinements", ct.c_int32), ("matvecs", ct.c_int32),
                ("rel_residual", ct.c_double),
                ("seconds", ct.c_double), ("status", ct.c_int32)]


class AdjointOpts(ct.Structure):
    _fields_ = [("lin_rel_tol", ct.c_double), ("picard_iterations", ct.c_int32),
                ("checkpoint_every", ct.c_int32), ("precision", ct.c_int32),
                ("gmres_restart", ct.c_int32), ("precond_sweeps", ct.c_int32),
                ("omega", ct.c_double), ("max_sweeps", ct.c_int32), ("precond_type", ct.c_int32)]


class AdjointStats(ct.Structure):
    _fields_ = [("picard_iterations", ct.c_int32), ("checkpoints", ct.c_int32),
                ("forward_sweeps", ct.c_int32), ("recompute_sweeps", ct.c_int32),
                ("adjoint_sweeps", ct.c_int32), ("adjoint_matvecs", ct.c_int32),
                ("max_forward_residual", ct.c_double), ("max_adjoint_residual", ct.c_double),
                ("forward_seconds", ct.c_double), ("backward_seconds", ct.c_double),
                ("status", ct.c_int32)]


class BatchOpts(ct.Structure):
    _fields_ = [("rel_tol", ct.c_double), ("precision", ct.c_int32), ("gmres_restart", ct.c_int32),
                ("precond_sweeps", ct.c_int32), ("omega", ct.c_double), ("max_cycles", ct.c_int32),
                ("use_graph", ct.c_int32), ("precond_type", ct.c_int32)]


class BatchStats(ct.Structure):
    _fields_ = [("cycles", ct.c_int32), ("sweeps", ct.c_int32), ("converged", ct.c_int32),
                ("max_rel_residual", ct.c_double), ("seconds", ct.c_double), ("status", ct.c_int32)]


class SweepTuning(ct.Structure):
    _fields_ = [("threads_per_block", ct.c_int32), ("schedule", ct.c_int32), ("loads", ct.c_int32),
                ("n_candidates", ct.c_int32), ("cand_threads", ct.c_int32 * 32),
                ("cand_schedule", ct.c_int32 * 32), ("cand_loads", ct.c_int32 * 32),
                ("cand_us_per_sweep", ct.c_double * 32)]

    def candidates(self):
        return [{"threads": self.cand_threads[i], "schedule": self.cand_schedule[i],
                 "loads": self.cand_loads[i], "us_per_sweep": round(self.cand_us_per_sweep[i], 2)}
                for i in range(self.n_candidates)]


_lib = None


def load():
    """Load libufe_dyn.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libufe_dyn.so not found at {LIB_PATH}; run "
                               "`python -m ufemism_amd.build` (or __graft_entry__.build())")
        # torch ships its own HIP runtime.  Both runtimes bind the HSA runtime by
        # soname, and torch's finds no device on a newer HSA runtime loaded before
        # it, so a process that may use torch loads torch first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ct.CDLL(LIB_PATH)
        lib.ufe_last_error.restype = ct.c_char_p
        lib.ufe_field_name.restype = ct.c_char_p
        lib.ufe_field_id.argtypes = [ct.c_char_p]
        lib.ufe_calculate_ice_thickness_change.argtypes = [_P, ct.c_double]
        lib.ufe_ssa_linear_solve.argtypes = [_P, _P, _P]
        lib.ufe_ssa_sor_solve_noalloc.argtypes = [_P, ct.c_double, ct.c_int32, ct.c_double, ct.c_int32, _P]
        lib.ufe_field_fill.argtypes = [_P, ct.c_int, ct.c_double]
        for name in EXPORTS:
            getattr(lib, name)  # raises AttributeError if a symbol is missing
        _lib = lib
    return _lib


def field_table():
    lib = load()
    out = {}
    for i in range(lib.ufe_field_count()):
        g, d = ct.c_int32(), ct.c_int32()
        lib.ufe_field_info(i, ct.byref(g), ct.byref(d))
        out[lib.ufe_field_name(i).decode()] = (i, g.value, d.value)
    return out


class UFEError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def _ptr(a):
    return None if a is None else ct.c_void_p(a.ctypes.data)


def params_struct(p: DynParams, precision=UFE_FP64, sweep_batch=0, linear_rel_tol=1e-6,
                  gmres_restart=0, precond_sweeps=0, precond_type=0) -> Params:
    s = Params()
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
    s.precision = int(precision)
    s.sweep_batch = int(sweep_batch)
    s.SSA_linear_rel_tol = float(linear_rel_tol)
    s.SSA_gmres_restart = int(gmres_restart)
    s.SSA_precond_sweeps = int(precond_sweeps)
    s.physical_properties = int(p.physical_properties)
    s.SSA_precond_type = int(precond_type)
    return s


class MeshArrays:
    """Primary mesh plus secondary arrays in reference layout (host)."""

    def __init__(self, mesh, nZ=15):
        self.mesh = mesh
        self.nV, self.nC_mem = mesh.nV, mesh.nC_mem
        self.nAc = int(mesh.nC.sum() // 2)
        self.nVAaAc = self.nV + self.nAc
        self.arr = F.alloc(F.MESH_SECONDARY, self.nV, self.nAc, self.nC_mem, nZ)
        self.colour_nV = [0] * 5

    def __getitem__(self, k):
        if k in self.arr:
            return self.arr[k]
        return getattr(self.mesh, k)

    def desc(self) -> MeshDesc:
        m, d = self.mesh, MeshDesc()
        d.nV, d.nTri, d.nC_mem, d.nAc = m.nV, m.nTri, m.nC_mem, self.nAc
        d.xmin, d.xmax, d.ymin, d.ymax = m.xmin, m.xmax, m.ymin, m.ymax
        for name in ("V", "nC", "C", "niTri", "iTri", "edge_index", "Tri", "A", "Cw"):
            a = self[name]
            assert a.flags.f_contiguous, name
            setattr(d, name, a.ctypes.data)
        for name in F.MESH_SECONDARY:
            setattr(d, name, self.arr[name].ctypes.data)
        for k in range(5):
            d.colour_nV[k] = self.colour_nV[k]
        self._keep = d
        return d


_contexts_opened = 0


def contexts_opened() -> int:
    """Contexts created in this process so far: after the first, the process
    holds a HIP runtime and must not fork workers."""
    return _contexts_opened


class Context:
    """One device context: uploaded mesh plus device-resident ice fields."""

    def __init__(self, device=0):
        global _contexts_opened
        self.lib = load()
        h = ct.c_void_p()
        _contexts_opened += 1
        r = self.lib.ufe_ctx_create(int(device), ct.byref(h))
        if r != UFE_OK:
            raise UFEError(r, "ufe_ctx_create failed (no HIP device?)")
        self.h = h
        self.fields = field_table()
        self.sizes = None

    def close(self):
        if self.h:
            self.lib.ufe_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, r, what):
        if r not in (UFE_OK,):
            raise UFEError(r, f"{what}: {self.lib.ufe_last_error(self.h).decode()}")
        return r

    def set_params(self, p: DynParams, precision=UFE_FP64, sweep_batch=0, **krylov):
        """krylov: linear_rel_tol, gmres_restart, precond_sweeps, precond_type (UFE_MIXED /
        UFE_FP64_KRYLOV; both take precond_type UFE_PRECOND_SOR or UFE_PRECOND_MULTIGRID)."""
        self._params = params_struct(p, precision, sweep_batch, **krylov)
        self._chk(self.lib.ufe_set_params(self.h, ct.byref(self._params)), "ufe_set_params")

    def mesh_geometry(self, mesh, compute_Tricc=True):
        """Tricc, A, Cw, TriA and R of a primary mesh on the GPU, as the reference
        computes them after creating or reading a mesh.  Returns a dict with
        those arrays, Tri (rotated where a circumcentre needed the retry) and the
        statistics."""
        nV, nT, nCm = mesh.nV, mesh.nTri, mesh.nC_mem
        out = {"Tri": np.array(mesh.Tri, dtype=np.int32, order="F"),
               "Tricc": (np.zeros((nT, 2), order="F") if compute_Tricc
                         else np.array(mesh.Tricc, dtype=np.float64, order="F")),
               "A": np.zeros(nV), "Cw": np.zeros((nV, nCm), order="F"), "TriA": np.zeros(nT),
               "R": np.zeros(nV)}
        keep = {}
        g = MeshGeometryDesc()
        g.nV, g.nTri, g.nC_mem = nV, nT, nCm
        g.xmin, g.xmax, g.ymin, g.ymax = mesh.xmin, mesh.xmax, mesh.ymin, mesh.ymax
        g.tol_dist = mesh.tol_dist
        for name, dt in (("V", np.float64), ("nC", np.int32), ("C", np.int32), ("niTri", np.int32),
                         ("iTri", np.int32), ("edge_index", np.int32), ("Tri_edge_index", np.int32)):
            keep[name] = np.asfortranarray(getattr(mesh, name), dtype=dt)
            setattr(g, name, keep[name].ctypes.data)
        for name in ("Tri", "Tricc", "A", "Cw", "TriA", "R"):
            setattr(g, name, out[name].ctypes.data)
        g.compute_Tricc = int(bool(compute_Tricc))
        st = MeshGeometryStats()
        self._chk(self.lib.ufe_mesh_geometry(self.h, ct.byref(g), ct.byref(st)), "ufe_mesh_geometry")
        out["stats"] = st
        return out

    def mesh_build(self, mesh, geometry=True) -> MeshArrays:
        """Secondary mesh data of a primary mesh.  With geometry (the default)
        the GPU first computes Tricc, A and Cw from the primary arrays, as the
        reference does after creating or reading a mesh; without it the mesh's
        own A and Cw are used."""
        ma = MeshArrays(mesh)
        if geometry:
            geo = self.mesh_geometry(mesh)
            for k in ("Tri", "Tricc", "A", "Cw", "TriA", "R"):
                ma.arr[k] = geo[k]
            ma.geometry_stats = geo["stats"]
        d = ma.desc()
        r = self.lib.ufe_mesh_build(self.h, ct.byref(d))
        ma.build_status = r
        if r != UFE_OK:
            raise UFEError(r, f"ufe_mesh_build: {self.lib.ufe_last_error(self.h).decode()}")
        ma.colour_nV = list(d.colour_nV)
        return ma

    def mesh_upload(self, ma: MeshArrays):
        d = ma.desc()
        self._chk(self.lib.ufe_mesh_upload(self.h, ct.byref(d)), "ufe_mesh_upload")
        s = (ct.c_int32 * 4)()
        self.lib.ufe_mesh_sizes(self.h, s)
        self.sizes = dict(nV=s[0], nAc=s[1], nVAaAc=s[2], nC_mem=s[3])
        self.nZ = self._params.nZ if hasattr(self, "_params") else 15

    def _shape(self, grid):
        s = self.sizes
        return {0: (s["nV"],), 1: (s["nV"], self.nZ), 2: (s["nV"], s["nC_mem"]),
                3: (s["nAc"],), 4: (s["nAc"], self.nZ), 5: (s["nVAaAc"],)}[grid]

    def set(self, name, value):
        fid, grid, dt = self.fields[name]
        a = np.asfortranarray(np.broadcast_to(np.asarray(value), self._shape(grid)),
                              dtype=np.float64 if dt == 0 else np.int32)
        self._chk(self.lib.ufe_field_set(self.h, fid, _ptr(a)), f"set {name}")

    def get(self, name):
        fid, grid, dt = self.fields[name]
        a = np.zeros(self._shape(grid), dtype=np.float64 if dt == 0 else np.int32, order="F")
        self._chk(self.lib.ufe_field_get(self.h, fid, _ptr(a)), f"get {name}")
        return a

    def update_general_ice_model_data(self):
        self._chk(self.lib.ufe_update_general_ice_model_data(self.h), "update_general_ice_model_data")

    def ice_physical_properties(self, mode=1):
        """ice_physical_properties on its own (mode 1 Arrhenius from Ti, 2 fixed benchmark values)."""
        self._chk(self.lib.ufe_ice_physical_properties(self.h, int(mode)), "ice_physical_properties")

    def solve_SIA(self):
        self._chk(self.lib.ufe_solve_SIA(self.h), "solve_SIA")

    def solve_SSA(self, allow=(UFE_OK, UFE_SOR_NOT_CONVERGED)) -> SSAStats:
        st = SSAStats()
        r = self.lib.ufe_solve_SSA(self.h, ct.byref(st))
        if r not in allow:
            raise UFEError(r, f"solve_SSA: {self.lib.ufe_last_error(self.h).decode()}")
        return st

    def calculate_ice_thickness_change(self, dt):
        self._chk(self.lib.ufe_calculate_ice_thickness_change(self.h, float(dt)),
                  "calculate_ice_thickness_change")

    # -- remapping from another context's mesh onto this one --------------------
    def remap_build(self, src: "Context", ma_src, ma_dst) -> RemapStats:
        """create_remapping_arrays from the mesh of src (ma_src) to this
        context's mesh (ma_dst); both built with their geometry."""
        gs, ks = geometry_desc(ma_src)
        gd, kd = geometry_desc(ma_dst)
        st = RemapStats()
        self._chk(self.lib.ufe_remap_build(self.h, src.h, ct.byref(gs), ct.byref(gd), ct.byref(st)),
                  "ufe_remap_build")
        del ks, kd
        return st

    def remap_field(self, src: "Context", method, d_src, d_dst, nz=1):
        """remap_field_dp on device pointers; method is a name or UFE_REMAP_*."""
        m = REMAP_METHODS.get(method, method)
        self._chk(self.lib.ufe_remap_field(self.h, src.h, ct.c_int(int(m)), ct.c_void_p(int(d_src)),
                                           ct.c_void_p(int(d_dst)), ct.c_int(int(nz))), "ufe_remap_field")

    def remap_ice_model(self, src: "Context"):
        self._chk(self.lib.ufe_remap_ice_model(self.h, src.h), "ufe_remap_ice_model")

    def remap_export(self):
        n = ct.c_int(0)
        self._chk(self.lib.ufe_remap_export(self.h, ct.byref(n), *([None] * 9)), "ufe_remap_export")
        nVd = self.sizes["nV"]
        out = {"nn": np.zeros(nVd, np.int32), "tvi": np.zeros((nVd, 3), np.int32, order="F"),
               "tw": np.zeros((nVd, 3), order="F"), "start": np.zeros(nVd, np.int32),
               "end": np.zeros(nVd, np.int32), "vi": np.zeros(n.value, np.int32), "w0": np.zeros(n.value),
               "w1x": np.zeros(n.value), "w1y": np.zeros(n.value)}
        ptrs = [_ptr(out[k]) for k in ("nn", "tvi", "tw", "start", "end", "vi", "w0", "w1x", "w1y")]
        self._chk(self.lib.ufe_remap_export(self.h, ct.byref(n), *ptrs), "ufe_remap_export")
        return out

    def solve_SIA_3D(self):
        self._chk(self.lib.ufe_solve_SIA_3D(self.h), "solve_SIA_3D")

    def critical_timesteps(self):
        """(dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min) after the 0.9 correction."""
        out = (ct.c_double * 3)()
        self._chk(self.lib.ufe_critical_timesteps(self.h, out), "critical_timesteps")
        return tuple(out)

    def upwind_derivatives_3D(self, U, V, d, ddx, ddy, nz):
        """get_upwind_derivative_vertex_3D on device pointers (nV x nz);
        returns the number of vertex-layer pairs without an upwind triangle."""
        missing = ct.c_int(0)
        self._chk(self.lib.ufe_op_upwind_derivatives_3D(
            self.h, ct.c_void_p(int(U)), ct.c_void_p(int(V)), ct.c_void_p(int(d)), ct.c_void_p(int(ddx)),
            ct.c_void_p(int(ddy)), ct.c_int(int(nz)), ct.byref(missing)), "upwind_derivatives_3D")
        return missing.value

    def ssa_prepare(self):
        self._chk(self.lib.ufe_ssa_prepare(self.h), "ssa_prepare")

    def ssa_linear_solve(self, rel_tol, max_sweeps, precision=UFE_FP64, omega=0.0,
                         check_interval=16, gmres=0, precond_sweeps=2,
                         orth_passes=2, method=0, full_sweeps=False,
                         neumann_every_sweep=False, precond_type=0) -> LinSolveStats:
        """method 0: GMRES(gmres) (plain SOR if gmres == 0); 1: BiCGStab.
        precond_type UFE_PRECOND_SOR: precond_sweeps SOR sweeps; UFE_PRECOND_MULTIGRID:
        an aggregation-multigrid K-cycle with precond_sweeps smoothing sweeps (mixed only)."""
        o = LinSolveOpts(float(rel_tol), int(max_sweeps), int(precision), float(omega),
                         int(check_interval), 0, int(gmres), int(precond_sweeps), int(orth_passes),
                         int(method), int(bool(full_sweeps)), int(bool(neumann_every_sweep)),
                         int(precond_type))
        st = LinSolveStats()
        self._chk(self.lib.ufe_ssa_linear_solve(self.h, ct.byref(o), ct.byref(st)),
                  "ssa_linear_solve")
        return st

    def ssa_sor_solve_noalloc(self, rel_tol, max_sweeps, omega=0.0, check_interval=16) -> LinSolveStats:
        """The plain FP64 SOR solve without any device allocation (scratch in a
        static device array of the library); the same results as
        ssa_linear_solve(rel_tol, max_sweeps, UFE_FP64, omega, check_interval)."""
        st = LinSolveStats()
        self._chk(self.lib.ufe_ssa_sor_solve_noalloc(self.h, float(rel_tol), int(max_sweeps), float(omega),
                                                     int(check_interval), ct.byref(st)), "ssa_sor_solve_noalloc")
        return st

    def multigrid_info(self) -> dict:
        """The multigrid hierarchy of the uploaded mesh (levels 0 before its first use)."""
        o = MultigridInfo()
        self._chk(self.lib.ufe_multigrid_info_get(self.h, ct.byref(o)), "multigrid_info")
        return o.as_dict()

    def set_device(self, name, ptr):
        """Copy a device buffer (reference layout) into a field."""
        fid = self.fields[name][0]
        self._chk(self.lib.ufe_field_set_device(self.h, fid, ct.c_void_p(int(ptr))), f"set_device {name}")

    def get_device(self, name, ptr):
        """Copy a field into a device buffer (reference layout)."""
        fid = self.fields[name][0]
        self._chk(self.lib.ufe_field_get_device(self.h, fid, ct.c_void_p(int(ptr))), f"get_device {name}")

    def adjoint_forward(self, picard_iterations, lin_rel_tol=1e-12, checkpoint_every=1,
                        precision=UFE_FP64, gmres_restart=40, precond_sweeps=5, omega=0.0,
                        max_sweeps=1_000_000, allow=(UFE_OK, UFE_SOR_NOT_CONVERGED),
                        precond_type=UFE_PRECOND_SOR) -> AdjointStats:
        o = AdjointOpts(float(lin_rel_tol), int(picard_iterations), int(checkpoint_every),
                        int(precision), int(gmres_restart), int(precond_sweeps), float(omega),
                        int(max_sweeps), int(precond_type))
        st = AdjointStats()
        r = self.lib.ufe_ssa_adjoint_forward(self.h, ct.byref(o), ct.byref(st))
        if r not in allow:
            raise UFEError(r, f"ssa_adjoint_forward: {self.lib.ufe_last_error(self.h).decode()}")
        if r == UFE_SOR_NOT_CONVERGED:
            warnings.warn(f"ssa_adjoint_forward: a linear solve stopped at relative residual "
                          f"{st.max_forward_residual:.2e} > {lin_rel_tol:.1e}", RuntimeWarning)
        return st

    def adjoint_backward(self, dJ_dU, dJ_dV, grad_log_tau_c, grad_A,
                         allow=(UFE_OK, UFE_SOR_NOT_CONVERGED)) -> AdjointStats:
        """Device pointers (ints) of nVAaAc doubles, reference AaAc order."""
        st = AdjointStats()
        r = self.lib.ufe_ssa_adjoint_backward(self.h, *(ct.c_void_p(int(x)) for x in
                                                        (dJ_dU, dJ_dV, grad_log_tau_c, grad_A)),
                                              ct.byref(st))
        if r not in allow:
            raise UFEError(r, f"ssa_adjoint_backward: {self.lib.ufe_last_error(self.h).decode()}")
        if r == UFE_SOR_NOT_CONVERGED:
            warnings.warn(f"ssa_adjoint_backward: an adjoint solve stopped at relative residual "
                          f"{st.max_adjoint_residual:.2e}", RuntimeWarning)
        return st

    def batch_define(self, nV, nAc):
        """Declare the member sizes of the union mesh uploaded next (empty: clear)."""
        nV = np.ascontiguousarray(nV, dtype=np.int32)
        nAc = np.ascontiguousarray(nAc, dtype=np.int32)
        self._batch_sizes = (nV, nAc)
        self._chk(self.lib.ufe_batch_define(self.h, int(len(nV)), _ptr(nV), _ptr(nAc)), "batch_define")

    def sweep_autotune(self, sweeps=20) -> SweepTuning:
        """Time every launch shape of the FP32 sweep on the uploaded mesh; keep the fastest."""
        t = SweepTuning()
        self._chk(self.lib.ufe_sweep_autotune(self.h, int(sweeps), ct.byref(t)), "sweep_autotune")
        return t

    def sweep_tuning_set(self, threads_per_block=256, schedule=0, loads=0):
        self._chk(self.lib.ufe_sweep_tuning_set(self.h, int(threads_per_block), int(schedule), int(loads)),
                  "sweep_tuning_set")

    def _batch_opts(self, rel_tol, precision, gmres_restart, precond_sweeps, omega, max_cycles, use_graph,
                    precond_type=UFE_PRECOND_SOR):
        return BatchOpts(float(rel_tol), int(precision), int(gmres_restart), int(precond_sweeps),
                         float(omega), int(max_cycles), int(bool(use_graph)), int(precond_type))

    def batch_linear_solve(self, rel_tol=1e-6, precision=UFE_MIXED, gmres_restart=40, precond_sweeps=5,
                           omega=0.0, max_cycles=0, use_graph=True, precond_type=UFE_PRECOND_SOR):
        o = self._batch_opts(rel_tol, precision, gmres_restart, precond_sweeps, omega, max_cycles, use_graph,
                             precond_type)
        st = BatchStats()
        rel = np.zeros(len(self._batch_sizes[0]))
        self._chk(self.lib.ufe_batch_linear_solve(self.h, ct.byref(o), _ptr(rel), ct.byref(st)),
                  "batch_linear_solve")
        return st, rel

    def batch_solve_SSA(self, picard_iterations, rel_tol=1e-6, precision=UFE_MIXED, gmres_restart=40,
                        precond_sweeps=5, omega=0.0, max_cycles=0, use_graph=True, precond_type=UFE_PRECOND_SOR):
        o = self._batch_opts(rel_tol, precision, gmres_restart, precond_sweeps, omega, max_cycles, use_graph,
                             precond_type)
        st = BatchStats()
        rel = np.zeros(len(self._batch_sizes[0]))
        self._chk(self.lib.ufe_batch_solve_SSA(self.h, int(picard_iterations), ct.byref(o), _ptr(rel),
                                               ct.byref(st)), "batch_solve_SSA")
        return st, rel

    def fill(self, name, value):
        fid = self.fields[name][0]
        self._chk(self.lib.ufe_field_fill(self.h, fid, float(value)), f"fill {name}")

    def profile_enable(self, on=True):
        self._chk(self.lib.ufe_profile_enable(self.h, int(bool(on))), "profile_enable")

    def profile_read(self) -> KernelProfile:
        p = KernelProfile()
        self._chk(self.lib.ufe_profile_read(self.h, ct.byref(p)), "profile_read")
        return p

    def sweep_bytes(self):
        b64, b32, nd = ct.c_double(), ct.c_double(), ct.c_double()
        rows = ct.c_int64()
        self._chk(self.lib.ufe_sweep_bytes(self.h, ct.byref(b64), ct.byref(b32), ct.byref(nd),
                                           ct.byref(rows)), "sweep_bytes")
        return dict(fp64=b64.value, fp32=b32.value, mean_degree=nd.value, rows=rows.value)

    def ssa_residual(self) -> float:
        v = ct.c_double()
        self._chk(self.lib.ufe_ssa_residual(self.h, ct.byref(v)), "ssa_residual")
        return v.value

    def op(self, name, *ptrs, nz=None):
        """Call ufe_op_<name> on device pointers (ints)."""
        args = [self.h] + [ct.c_void_p(int(p)) for p in ptrs]
        if nz is not None:
            args.append(ct.c_int(int(nz)))
        self._chk(getattr(self.lib, "ufe_op_" + name)(*args), name)

    def synchronize(self):
        self._chk(self.lib.ufe_synchronize(self.h), "synchronize")
