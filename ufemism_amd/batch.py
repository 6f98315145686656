"""Batched ensembles: many independent SSA problems solved together on one GPU.

Each member is an ordinary workload (mesh + ice state).  Its secondary mesh
data (Ac mesh, AaAc mesh, neighbour functions, five-colouring;
src/mesh_ArakawaC_module.f90:24-540, src/mesh_derivatives_module.f90:14-313,
src/mesh_five_colour_module.f90:18-735) is built on its own, then the members
are joined into one union mesh in the reference layout:

* Aa vertices member-major (member 0's vertices, member 1's, ...), then the
  Ac points member-major; every index array is shifted accordingly;
* the colour lists of all members concatenated per colour;
* each member keeps its four corner vertices (edge_index 2/4/6/8).

The union is uploaded after ``ufe_batch_define`` has declared the member
sizes, so the library keeps each member's rows together inside every colour
and gives each member its own reductions and Krylov coefficients.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import abi, fields as F, geometry
from .dynamics import initial_fields


def _member(spec):
    kind, size, s, A, friction_sigma, picard = spec
    if kind == 0:
        wl = geometry.dome(size, seed=s, picard=picard)
    else:
        wl = geometry.marine(size, seed=s, picard=picard)
    wl.A_flow = A
    wl.friction_sigma = friction_sigma
    wl.friction_seed = s + 77
    return wl


def ensemble(n=64, nv_min=5000, nv_max=100_000, seed=2024, friction_sigma=0.5, picard=5, workers=1):
    """Seeded members: sizes log-uniform in [nv_min, nv_max], dome or marine
    geometry by a seeded coin, A uniform in [1e-17, 1e-16] Pa^-3 yr^-1 and a
    log-normal friction field (sigma friction_sigma).  With workers > 1 the
    meshes are generated in forked processes: call it before the process
    opens a GPU context."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nv_min), np.log(nv_max), n)).astype(int)
    kinds = rng.integers(0, 2, n)
    A = rng.uniform(1e-17, 1e-16, n)
    specs = [(int(kinds[b]), int(sizes[b]), int(seed * 1000 + b), float(A[b]), friction_sigma, picard)
             for b in range(n)]
    if workers > 1:
        import concurrent.futures
        import multiprocessing
        with concurrent.futures.ProcessPoolExecutor(
                max_workers=workers, mp_context=multiprocessing.get_context("fork")) as pool:
            return list(pool.map(_member, specs))
    return [_member(sp) for sp in specs]


class _UnionMesh:
    """Primary fields of the union, with the attribute names of meshgen.Mesh."""


@dataclasses.dataclass
class Member:
    wl: object
    nV: int
    nAc: int
    vbase: int
    acbase: int


def join_meshes(built, members, nZ=15):
    """Union of the members' meshes in the reference layout (see the module
    docstring).  `built[b]` holds member b's primary mesh (`.mesh`), its
    secondary arrays (`[name]`) and `.colour_nV`; `members[b]` its offsets."""
    ms = members
    nV = sum(m.nV for m in ms)
    nAc = sum(m.nAc for m in ms)
    nA = nV + nAc
    nCm = built[0].mesh.nC_mem
    um = _UnionMesh()
    um.nV, um.nC_mem = nV, nCm
    um.nTri = int(sum(ma.mesh.nTri for ma in built))
    um.xmin = min(ma.mesh.xmin for ma in built)
    um.xmax = max(ma.mesh.xmax for ma in built)
    um.ymin = min(ma.mesh.ymin for ma in built)
    um.ymax = max(ma.mesh.ymax for ma in built)

    def cat(get, shift=None):
        parts = []
        for m, ma in zip(ms, built):
            a = np.asarray(get(ma))
            if shift is not None:
                a = np.where(a > 0, a + shift(m), a)
            parts.append(a)
        return np.asfortranarray(np.concatenate(parts, axis=0))

    tb = np.cumsum([0] + [ma.mesh.nTri for ma in built])
    um.V = cat(lambda ma: ma.mesh.V)
    um.nC = cat(lambda ma: ma.mesh.nC)
    um.C = cat(lambda ma: ma.mesh.C, lambda m: m.vbase)
    um.niTri = cat(lambda ma: ma.mesh.niTri)
    tri_base = {id(m): int(tb[i]) for i, m in enumerate(ms)}
    um.iTri = cat(lambda ma: ma.mesh.iTri, lambda m: tri_base[id(m)])
    um.edge_index = cat(lambda ma: ma.mesh.edge_index)
    um.Tri = cat(lambda ma: ma["Tri"], lambda m: m.vbase)
    um.A = cat(lambda ma: ma["A"])
    um.Cw = cat(lambda ma: ma["Cw"])
    ua = abi.MeshArrays.__new__(abi.MeshArrays)
    ua.mesh = um
    ua.nV, ua.nC_mem, ua.nAc, ua.nVAaAc = nV, nCm, nAc, nA
    ua.arr = F.alloc(F.MESH_SECONDARY, nV, nAc, nCm, nZ)
    A = ua.arr
    for k in ("Nx", "Ny", "Nxx", "Nxy", "Nyy", "Cw"):
        if k in A:
            A[k][...] = cat(lambda ma, k=k: ma[k])
    A["iAci"][...] = cat(lambda ma: ma["iAci"], lambda m: m.acbase)
    A["Aci"][...] = cat(lambda ma: ma["Aci"], lambda m: m.vbase)
    # boundary Ac rows carry vertex 1 with a zero weight as their 4th entry
    # (src/mesh_ArakawaC_module.f90:165-167); keep it as the reference writes it
    bnd = cat(lambda ma: ma["edge_index_Ac"]) > 0
    A["Aci"][bnd, 3] = cat(lambda ma: ma["Aci"][:, 3])[bnd]
    for k in ("VAc", "Nx_Ac", "Ny_Ac", "No_Ac", "Np_Ac", "edge_index_Ac"):
        A[k][...] = cat(lambda ma, k=k: ma[k])
    # AaAc arrays: rows move to their union positions, indices are remapped
    colour_lists = [[] for _ in range(5)]
    for m, ma in zip(ms, built):
        g = aaac_map(m, nV)
        look = np.concatenate([[0], g + 1])          # 1-based member index -> 1-based union index
        for k in ("VAaAc", "nCAaAc", "Nx_AaAc", "Ny_AaAc", "Nxx_AaAc", "Nxy_AaAc", "Nyy_AaAc", "colour"):
            A[k][g, ...] = ma[k]
        A["CAaAc"][g, :] = look[ma["CAaAc"]]
        for c in range(5):
            colour_lists[c].append(look[ma["colour_vi"][: ma.colour_nV[c], c]])
    for c in range(5):
        lst = np.concatenate(colour_lists[c])
        A["colour_vi"][: len(lst), c] = lst
    ua.colour_nV = [int(sum(ma.colour_nV[c] for ma in built)) for c in range(5)]
    return ua


def aaac_map(m: Member, nV_union: int):
    """Union 0-based AaAc index of each of member m's AaAc points."""
    loc = np.arange(m.nV + m.nAc)
    return np.where(loc < m.nV, loc + m.vbase, nV_union + m.acbase + (loc - m.nV))


def members_of(built):
    out, vb, ab = [], 0, 0
    for ma in built:
        nV = ma.mesh.nV
        nAc = int(ma.nAc)
        out.append(Member(None, nV, nAc, vb, ab))
        vb += nV
        ab += nAc
    return out


class Batch:
    """The union of the members' meshes, uploaded to one context."""

    def __init__(self, ctx: abi.Context, workloads, nZ=15):
        self.ctx = ctx
        self.workloads = list(workloads)
        p0 = self.workloads[0].params
        ctx.set_params(p0, abi.UFE_FP64)
        built = [ctx.mesh_build(wl.mesh) for wl in self.workloads]
        self.members = members_of(built)
        for m, wl in zip(self.members, self.workloads):
            m.wl = wl
        self.nV = sum(m.nV for m in self.members)
        self.nAc = sum(m.nAc for m in self.members)
        self.nA = self.nV + self.nAc
        self.arrays = join_meshes(built, self.members, nZ)
        nVs = np.array([m.nV for m in self.members], np.int32)
        nAcs = np.array([m.nAc for m in self.members], np.int32)
        ctx.batch_define(nVs, nAcs)
        ctx.mesh_upload(self.arrays)
        self.init = self._join_fields(built, nZ)
        self.reset()

    # -- union mesh -------------------------------------------------------------
    def aaac_map(self, m: Member):
        return aaac_map(m, self.nV)

    def _join_fields(self, built, nZ):
        """Initial fields of the union (reference layout), keyed by field name."""
        per = [initial_fields(m.wl, ma["VAaAc"], nZ) for m, ma in zip(self.members, built)]
        out = {}
        for k in ("Hi", "Hb", "SL", "A_flow_mean", "SMB_year", "BMB"):
            out[k] = np.concatenate([np.broadcast_to(f[k], (m.nV,)) for f, m in zip(per, self.members)])
        out["A_flow"] = np.asfortranarray(np.concatenate([np.broadcast_to(f["A_flow"], (m.nV, nZ))
                                                          for f, m in zip(per, self.members)], 0))
        out["A_flow_Ac"] = np.asfortranarray(np.concatenate([np.full((m.nAc, nZ), f["A_flow_Ac"])
                                                             for f, m in zip(per, self.members)], 0))
        out["A_flow_mean_Ac"] = np.concatenate([np.full(m.nAc, f["A_flow_mean_Ac"])
                                                for f, m in zip(per, self.members)])
        tau = np.zeros(self.nA)
        for f, m in zip(per, self.members):
            tau[self.aaac_map(m)] = f["tau_c_override_AaAc"]
        out["tau_c_override_AaAc"] = tau
        return out

    # -- state ------------------------------------------------------------------
    def reset(self):
        for k, v in self.init.items():
            self.ctx.set(k, v)
        for k in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
            self.ctx.fill(k, 0.0)

    def member_values(self, name):
        """Per-member slices of a union field (Aa, Ac or AaAc grid)."""
        a = self.ctx.get(name)
        grid = self.ctx.fields[name][1]
        out = []
        for m in self.members:
            if grid in (0, 1, 2):
                out.append(a[m.vbase:m.vbase + m.nV])
            elif grid in (3, 4):
                out.append(a[m.acbase:m.acbase + m.nAc])
            else:
                out.append(a[self.aaac_map(m)])
        return out
