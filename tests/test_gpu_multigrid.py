"""The aggregation multigrid preconditioner of the mixed-precision Krylov
solves (UFE_PRECOND_MULTIGRID) on the 20k marine system:

* GMRES-IR and BiCGStab-IR with it reach the FP64 tolerance and land on the
  FP64 Krylov solution, in fewer Krylov iterations than with SOR sweeps;
* the hierarchy coarsens the SOR rows down to a level that is solved by a
  dense inverse;
* a second solve of the same system gives the same bits (no atomics, fixed
  reduction order);
* solve_SSA's Picard loop with SSA_precond_type = multigrid agrees with the
  FP64 Krylov step, on the marine sheet and on a grounded dome;
* FP64 Krylov vectors around the FP32 cycle reach 1e-11; plain SOR with the
  multigrid preconditioner is refused;
* with the grounding-line flux on, the Krylov modes keep the prescribed rows
  and agree with each other and with the reference SOR step.
"""
import dataclasses

import numpy as np
import pytest

from ufemism_amd import abi, geometry
from ufemism_amd.dynamics import Step

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _linear_system(ctx, n=20000, seed=5, picard=4):
    wl = geometry.marine(n, seed=seed, picard=picard)
    step = Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    ctx.solve_SSA()
    ctx.ssa_prepare()
    return wl, step


def _solve(ctx, tol, **kw):
    ctx.fill("U_SSA_AaAc", 0.0)
    ctx.fill("V_SSA_AaAc", 0.0)
    st = ctx.ssa_linear_solve(tol, 1_000_000, omega=1.3, orth_passes=1, **kw)
    uv = np.concatenate([ctx.get("U_SSA_AaAc"), ctx.get("V_SSA_AaAc")])
    return st, uv


@pytest.fixture(scope="module")
def system(ctx):
    print("[multigrid] marine20k linearised system", flush=True)
    _linear_system(ctx)
    st, ref = _solve(ctx, 1e-11, precision=abi.UFE_FP64, gmres=40, precond_sweeps=5)
    assert st.status == 0, st.rel_residual
    return ref


def test_multigrid_gmres_matches_fp64(ctx, system):
    ref = system
    st_s, uv_s = _solve(ctx, 1e-9, precision=abi.UFE_MIXED, gmres=40, precond_sweeps=5)
    st_m, uv_m = _solve(ctx, 1e-9, precision=abi.UFE_MIXED, gmres=40, precond_sweeps=2,
                        precond_type=abi.UFE_PRECOND_MULTIGRID)
    print(f"  sor: matvecs {st_s.matvecs} refinements {st_s.refinements}; multigrid: matvecs {st_m.matvecs} "
          f"refinements {st_m.refinements} rel {st_m.rel_residual:.1e}", flush=True)
    assert st_m.status == 0 and st_m.rel_residual <= 1e-9
    assert abs(ctx.ssa_residual() - st_m.rel_residual) <= 1e-12
    e = _rel(uv_m, ref)
    print(f"  multigrid vs fp64 rel L2 {e:.2e} (sor {_rel(uv_s, ref):.2e})", flush=True)
    assert e <= 1e-6
    assert st_m.matvecs * 2 <= st_s.matvecs
    info = ctx.multigrid_info()
    print(f"  hierarchy {info}", flush=True)
    pts = info["points"]
    assert info["levels"] >= 2 and info["direct"]
    assert all(pts[k + 1] < pts[k] for k in range(len(pts) - 1))
    assert pts[-1] <= 64 and not info["kstep"][-1]


def test_multigrid_bicgstab_matches_fp64(ctx, system):
    ref = system
    st, uv = _solve(ctx, 1e-9, precision=abi.UFE_MIXED, gmres=8, precond_sweeps=2, method=1,
                    precond_type=abi.UFE_PRECOND_MULTIGRID)
    e = _rel(uv, ref)
    print(f"  bicgstab+multigrid: matvecs {st.matvecs} rel {st.rel_residual:.1e}, vs fp64 {e:.2e}", flush=True)
    assert st.status == 0 and st.rel_residual <= 1e-9
    assert e <= 1e-6


def test_multigrid_solve_is_repeatable(ctx, system):
    a = _solve(ctx, 1e-8, precision=abi.UFE_MIXED, gmres=40, precond_sweeps=2,
               precond_type=abi.UFE_PRECOND_MULTIGRID)
    b = _solve(ctx, 1e-8, precision=abi.UFE_MIXED, gmres=40, precond_sweeps=2,
               precond_type=abi.UFE_PRECOND_MULTIGRID)
    assert a[0].matvecs == b[0].matvecs
    assert np.array_equal(a[1], b[1])


def test_multigrid_needs_a_krylov_method(ctx, system):
    for prec in (abi.UFE_FP64, abi.UFE_MIXED):
        with pytest.raises(abi.UFEError):
            ctx.ssa_linear_solve(1e-6, 100, precision=prec, gmres=0, precond_type=abi.UFE_PRECOND_MULTIGRID)


def test_multigrid_with_fp64_krylov(ctx, system):
    """FP64 Krylov vectors around the FP32 cycle reach 1e-11 directly."""
    ref = system
    st, uv = _solve(ctx, 1e-11, precision=abi.UFE_FP64, gmres=20, precond_sweeps=2,
                    precond_type=abi.UFE_PRECOND_MULTIGRID)
    print(f"  fp64 krylov + multigrid: matvecs {st.matvecs} refinements {st.refinements} rel {st.rel_residual:.1e}",
          flush=True)
    assert st.status == 0 and st.rel_residual <= 1e-11
    assert _rel(uv, ref) <= 1e-6          # two solutions to a 1e-11 residual: within the conditioning


def test_picard_loop_with_multigrid():
    """solve_SSA with multigrid-preconditioned mixed solves agrees with the FP64
    Krylov step, both solving every linearised system to 1e-10."""
    print("[multigrid ssa] marine20k, mixed+multigrid vs fp64 Krylov at 1e-10", flush=True)
    ctx = abi.Context(0)
    try:
        wl = geometry.marine(20000, seed=5, picard=10)
        step = Step(ctx, wl)
        out = {}
        for mode, pc in ((abi.UFE_FP64_KRYLOV, abi.UFE_PRECOND_SOR), (abi.UFE_MIXED, abi.UFE_PRECOND_MULTIGRID)):
            step.reset()
            for f in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
                ctx.fill(f, 0.0)
            ctx.set_params(wl.params, mode, linear_rel_tol=1e-10, precond_type=pc,
                           precond_sweeps=2 if pc else 5)
            r = step.run()
            assert r["stats"].status == 0, r["stats"].status
            out[mode] = r
            print(f"  mode {mode} precond {pc}: picard {r['picard']} sweeps {r['sweeps']}", flush=True)
        a, b = out[abi.UFE_FP64_KRYLOV], out[abi.UFE_MIXED]
        for k in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Hi"):
            e = _rel(b[k], a[k])
            print(f"  {k}: rel L2 {e:.2e}", flush=True)
            assert e <= 1e-7, (k, e)
    finally:
        ctx.close()


def test_multigrid_on_a_grounded_dome():
    """A fully grounded dome with ice-free margins (large sliding terms around
    it): the multigrid-preconditioned Picard loop agrees with the FP64 Krylov
    step."""
    print("[multigrid ssa] dome8k, mixed+multigrid vs fp64 Krylov at 1e-10", flush=True)
    ctx = abi.Context(0)
    try:
        wl = geometry.dome(8000, seed=1, picard=10)
        step = Step(ctx, wl)
        out = {}
        for mode, pc in ((abi.UFE_FP64_KRYLOV, abi.UFE_PRECOND_SOR), (abi.UFE_MIXED, abi.UFE_PRECOND_MULTIGRID)):
            step.reset()
            for f in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
                ctx.fill(f, 0.0)
            ctx.set_params(wl.params, mode, linear_rel_tol=1e-10, precond_type=pc)
            r = step.run()
            assert r["stats"].status == 0, r["stats"].status
            out[mode] = r
            print(f"  mode {mode} precond {pc}: picard {r['picard']} sweeps {r['sweeps']}", flush=True)
        info = ctx.multigrid_info()
        assert info["levels"] >= 2 and info["direct"]
        a, b = out[abi.UFE_FP64_KRYLOV], out[abi.UFE_MIXED]
        for k in ("U_SSA", "V_SSA", "Hi"):
            e = _rel(b[k], a[k])
            print(f"  {k}: rel L2 {e:.2e}", flush=True)
            assert e <= 1e-7, (k, e)
    finally:
        ctx.close()


def test_gl_flux_with_the_krylov_modes():
    """use_analytical_GL_flux with the Krylov solves of solve_SSA: the rows that
    calculate_GL_flux prescribes keep those values bit for bit (the reference
    SOR skips them, src/ice_dynamics_module.f90:627-631); FP64 Krylov, mixed
    GMRES-IR with SOR sweeps and mixed GMRES-IR with the multigrid agree to
    1e-7; and all three agree with the reference SOR step run to
    max|dU| < 1e-4 m/yr."""
    print("[gl krylov] marine20k with the GL flux: reference SOR, fp64 Krylov, mixed sor, mixed multigrid",
          flush=True)
    ctx = abi.Context(0)
    try:
        wl = geometry.marine(20000, seed=5, picard=10)
        wl.A_flow = 1e-18
        wl.params.use_analytical_GL_flux = True
        step = Step(ctx, wl)
        tight = dataclasses.replace(wl.params, SSA_max_residual_UV=1e-4, SSA_max_inner_loops=1_000_000)
        runs = (("sor", tight, abi.UFE_FP64, abi.UFE_PRECOND_SOR),
                ("fp64_krylov", wl.params, abi.UFE_FP64_KRYLOV, abi.UFE_PRECOND_SOR),
                ("mixed_sor", wl.params, abi.UFE_MIXED, abi.UFE_PRECOND_SOR),
                ("mixed_multigrid", wl.params, abi.UFE_MIXED, abi.UFE_PRECOND_MULTIGRID))
        out, gl = {}, {}
        for name, params, mode, pc in runs:
            step.reset()
            ctx.set_params(params, mode, linear_rel_tol=1e-10, precond_type=pc)
            r = step.run()
            assert r["stats"].status == 0, (name, r["stats"].status)
            mask = ctx.get("mask_gl_Ac") == 1
            out[name] = r
            gl[name] = (mask, r["Ux_SSA_Ac"][mask], r["Uy_SSA_Ac"][mask])
            print(f"  {name}: picard {r['picard']} sweeps {sum(r['sweeps'])} GL rows {int(mask.sum())}", flush=True)
        m0, ux0, uy0 = gl["sor"]
        assert m0.sum() > 20
        for name in gl:
            m, ux, uy = gl[name]
            assert np.array_equal(m, m0) and np.array_equal(ux, ux0) and np.array_equal(uy, uy0), name
        a = out["fp64_krylov"]
        for name in ("mixed_sor", "mixed_multigrid"):
            for k in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Hi"):
                e = _rel(out[name][k], a[k])
                print(f"  {name} {k}: vs fp64 Krylov rel L2 {e:.2e}", flush=True)
                assert e <= 1e-7, (name, k, e)
        for k in ("U_SSA", "V_SSA", "Hi"):
            e = _rel(a[k], out["sor"][k])
            print(f"  fp64 Krylov {k}: vs reference SOR rel L2 {e:.2e}", flush=True)
            assert e <= 1e-3, (k, e)
    finally:
        ctx.close()


@pytest.mark.parametrize("n", [150, 300])
def test_multigrid_on_tiny_meshes(n):
    """Meshes so small that level 1 is (nearly) the dense coarsest level: the
    two-level and three-level hierarchies solve the Picard loop like FP64 SOR
    preconditioning does."""
    ctx = abi.Context(0)
    try:
        wl = geometry.dome(n, seed=2, picard=4)
        step = Step(ctx, wl)
        out = {}
        for mode, pc in ((abi.UFE_FP64_KRYLOV, abi.UFE_PRECOND_SOR), (abi.UFE_MIXED, abi.UFE_PRECOND_MULTIGRID)):
            step.reset()
            for f in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
                ctx.fill(f, 0.0)
            ctx.set_params(wl.params, mode, linear_rel_tol=1e-10, precond_type=pc)
            r = step.run()
            assert r["stats"].status == 0, r["stats"].status
            out[mode] = r
        info = ctx.multigrid_info()
        print(f"  dome nV={wl.mesh.nV}: hierarchy {info['points']} direct {info['direct']}", flush=True)
        assert info["direct"] and info["levels"] >= 2
        a, b = out[abi.UFE_FP64_KRYLOV], out[abi.UFE_MIXED]
        for k in ("U_SSA", "V_SSA", "Hi"):
            assert _rel(b[k], a[k]) <= 1e-7, (k, _rel(b[k], a[k]))
    finally:
        ctx.close()


@pytest.mark.parametrize("images", ["1", "0"])
def test_small_level_kernel_is_bit_identical(images):
    """The levels below the last K-step level solved by one workgroup per visit
    (k_amg_tail), with their constant arrays staged in LDS from the set-up's
    images or read from global memory (UFE_AMG_TAIL_IMAGES=0), give the same
    bits as one kernel per step (UFE_AMG_TAIL=0 at set-up): the forward mixed
    solve and both directions of the multigrid adjoint, whose backward cycles
    the transposed hierarchy."""
    import os
    import torch
    from ufemism_amd.adjoint import AdjointOptions, SSAProblem, ssa_velocities
    out = {}
    os.environ["UFE_AMG_TAIL_IMAGES"] = images
    try:
        for flag in ("1", "0"):
            os.environ["UFE_AMG_TAIL"] = flag
            ctx = abi.Context(0)
            try:
                wl, step = _linear_system(ctx)
                st, uv = _solve(ctx, 1e-9, precision=abi.UFE_MIXED, gmres=20, precond_sweeps=2,
                                precond_type=abi.UFE_PRECOND_MULTIGRID)
                info = ctx.multigrid_info()
                print(f"  UFE_AMG_TAIL={flag}: hierarchy {info['points']} tail level {info['tail_level']} "
                      f"matvecs {st.matvecs}", flush=True)
                assert st.status == 0
                assert (info["tail_level"] > 0) == (flag == "1")
                step.reset()
                ctx.update_general_ice_model_data()
                prob = SSAProblem(ctx, AdjointOptions(picard_iterations=3, lin_rel_tol=1e-10, gmres_restart=20,
                                                      precond_sweeps=2, precond_type=abi.UFE_PRECOND_MULTIGRID))
                lt = torch.log(torch.tensor(wl.tau_c_override(step.ma["VAaAc"]), dtype=torch.float64,
                                            device=prob.dev)).requires_grad_(True)
                A0 = torch.full((prob.nA,), wl.A_flow, dtype=torch.float64, device=prob.dev)
                U, V = ssa_velocities(lt, A0, prob)
                (0.5 * (U ** 2 + V ** 2).sum()).backward()
                out[flag] = (st.matvecs, uv, U.detach().cpu().numpy(), lt.grad.cpu().numpy())
            finally:
                ctx.close()
    finally:
        os.environ.pop("UFE_AMG_TAIL", None)
        os.environ.pop("UFE_AMG_TAIL_IMAGES", None)
    a, b = out["1"], out["0"]
    assert a[0] == b[0]
    for k in (1, 2, 3):
        assert np.array_equal(a[k], b[k]), k


def test_fused_cycle_kernels_are_bit_identical():
    """The cycle steps of the coarse levels as two kernels with block halos
    (k_amg_fuse_pre, k_amg_fuse_post) give the same bits as one kernel per
    step (UFE_AMG_FUSE=0 at set-up): for 1, 2 and 3 smoothing sweeps, with
    FP32 and FP64 Krylov vectors, and in both directions of the multigrid
    adjoint (the backward cycles the transposed hierarchy)."""
    import os
    import torch
    from ufemism_amd.adjoint import AdjointOptions, SSAProblem, ssa_velocities
    out = {}
    try:
        for flag in ("1", "0"):
            os.environ["UFE_AMG_FUSE"] = flag
            ctx = abi.Context(0)
            try:
                wl, step = _linear_system(ctx)
                res = []
                for prec, nu in ((abi.UFE_MIXED, 1), (abi.UFE_MIXED, 2), (abi.UFE_MIXED, 3), (abi.UFE_FP64, 2)):
                    st, uv = _solve(ctx, 1e-9, precision=prec, gmres=20, precond_sweeps=nu,
                                    precond_type=abi.UFE_PRECOND_MULTIGRID)
                    assert st.status == 0
                    res.append((st.matvecs, uv))
                info = ctx.multigrid_info()
                print(f"  UFE_AMG_FUSE={flag}: hierarchy {info['points']} fused {info['fused']} "
                      f"lists {info['fuse_list_max']} matvecs {[m for m, _ in res]}", flush=True)
                assert any(info["fused"]) == (flag == "1")
                step.reset()
                ctx.update_general_ice_model_data()
                prob = SSAProblem(ctx, AdjointOptions(picard_iterations=3, lin_rel_tol=1e-10, gmres_restart=20,
                                                      precond_sweeps=2, precond_type=abi.UFE_PRECOND_MULTIGRID))
                lt = torch.log(torch.tensor(wl.tau_c_override(step.ma["VAaAc"]), dtype=torch.float64,
                                            device=prob.dev)).requires_grad_(True)
                A0 = torch.full((prob.nA,), wl.A_flow, dtype=torch.float64, device=prob.dev)
                U, V = ssa_velocities(lt, A0, prob)
                (0.5 * (U ** 2 + V ** 2).sum()).backward()
                res.append((prob.backward_stats.adjoint_matvecs, np.concatenate([U.detach().cpu().numpy(),
                                                                                 lt.grad.cpu().numpy()])))
                out[flag] = res
            finally:
                ctx.close()
    finally:
        os.environ.pop("UFE_AMG_FUSE", None)
    for (ma, ua), (mb, ub) in zip(out["1"], out["0"]):
        assert ma == mb
        assert np.array_equal(ua, ub)


@pytest.mark.parametrize("depth", [1, 2])
def test_fused_cycle_kernels_with_shallow_halos(depth):
    """Halos of depth 1 or 2 (UFE_AMG_FUSE_DEPTH, the fallback for lists too
    long for LDS): the fused kernels run for up to that many sweeps, the K-step
    dots go through k_amg_dot_blocks, and the results equal the per-step
    kernels bit for bit (UFE_AMG_FUSE=0 with the same blocks)."""
    import os
    out = {}
    try:
        os.environ["UFE_AMG_FUSE_DEPTH"] = str(depth)
        for flag in ("1", "0"):
            os.environ["UFE_AMG_FUSE"] = flag
            ctx = abi.Context(0)
            try:
                _linear_system(ctx)
                res = []
                for nu in (1, 2):
                    st, uv = _solve(ctx, 1e-9, precision=abi.UFE_MIXED, gmres=20, precond_sweeps=nu,
                                    precond_type=abi.UFE_PRECOND_MULTIGRID)
                    assert st.status == 0
                    res.append((st.matvecs, uv))
                info = ctx.multigrid_info()
                print(f"  depth {depth} UFE_AMG_FUSE={flag}: fused {info['fused']} matvecs {[m for m, _ in res]}",
                      flush=True)
                if flag == "1":
                    assert max(info["fused"]) == depth
                out[flag] = res
            finally:
                ctx.close()
    finally:
        os.environ.pop("UFE_AMG_FUSE", None)
        os.environ.pop("UFE_AMG_FUSE_DEPTH", None)
    for (ma, ua), (mb, ub) in zip(out["1"], out["0"]):
        assert ma == mb
        assert np.array_equal(ua, ub)
