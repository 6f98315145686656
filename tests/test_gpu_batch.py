"""Batched ensembles on the GPU: every member of a batch solve must match the
same member solved alone, mixed precision must match FP64, hipGraph replays
must equal eager launches bit for bit, and the autotuned launch shape must
not change a single bit either (the shape only reorders independent rows)."""
import numpy as np
import pytest

from ufemism_amd import abi, batch
from ufemism_amd.dynamics import Step

pytestmark = pytest.mark.gpu

PICARD = 3
TOL = 1e-10


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def ens():
    print("[batch] building 6 members", flush=True)
    wls = batch.ensemble(n=6, nv_min=1500, nv_max=6000, seed=7, picard=PICARD)
    for wl in wls:
        wl.params.SSA_RN_tol = 0.0           # exactly PICARD iterations everywhere
    ctx = abi.Context(0)
    B = batch.Batch(ctx, wls)
    yield ctx, B, wls
    ctx.close()


def _solve(ctx, B, **kw):
    B.reset()
    ctx.update_general_ice_model_data()
    st, rel = ctx.batch_solve_SSA(PICARD, rel_tol=TOL, **kw)
    assert st.status == 0 and st.converged == len(B.members), (st.status, st.converged, rel)
    return B.member_values("U_SSA_AaAc"), B.member_values("V_SSA_AaAc"), st


def test_batch_members_match_single_solves(ens):
    ctx, B, wls = ens
    U, V, st = _solve(ctx, B, precision=abi.UFE_FP64)
    print(f"  fp64 batch: cycles {st.cycles} sweeps {st.sweeps} {st.seconds:.3f}s", flush=True)
    one = abi.Context(0)
    try:
        for b, wl in enumerate(wls):
            step = Step(one, wl)
            one.set_params(wl.params, abi.UFE_FP64_KRYLOV, linear_rel_tol=TOL)
            one.update_general_ice_model_data()
            s = one.solve_SSA()
            assert s.picard_iterations == PICARD
            u, v = one.get("U_SSA_AaAc"), one.get("V_SSA_AaAc")
            e = max(_rel(U[b], u), _rel(V[b], v))
            print(f"  member {b} nV={wl.mesh.nV}: batch vs alone rel L2 {e:.2e}", flush=True)
            assert e <= 1e-7, (b, e)
    finally:
        one.close()


def test_batch_mixed_matches_fp64(ens):
    ctx, B, wls = ens
    U0, V0, _ = _solve(ctx, B, precision=abi.UFE_FP64)
    U1, V1, st = _solve(ctx, B, precision=abi.UFE_MIXED)
    print(f"  mixed batch: cycles {st.cycles} sweeps {st.sweeps} {st.seconds:.3f}s", flush=True)
    for b in range(len(wls)):
        e = max(_rel(U1[b], U0[b]), _rel(V1[b], V0[b]))
        print(f"  member {b}: mixed vs fp64 rel L2 {e:.2e}", flush=True)
        assert e <= 1e-7, (b, e)


def test_graph_replays_equal_eager_launches(ens):
    ctx, B, wls = ens
    Ug, Vg, sg = _solve(ctx, B, precision=abi.UFE_MIXED, use_graph=True)
    Ue, Ve, se = _solve(ctx, B, precision=abi.UFE_MIXED, use_graph=False)
    assert sg.cycles == se.cycles and sg.cycles > PICARD
    for b in range(len(wls)):
        assert np.array_equal(Ug[b], Ue[b]) and np.array_equal(Vg[b], Ve[b]), b


def test_autotuned_shape_changes_no_bits(ens):
    ctx, B, wls = ens
    U0, V0, _ = _solve(ctx, B, precision=abi.UFE_MIXED)
    B.reset()
    ctx.update_general_ice_model_data()
    ctx.ssa_prepare()
    t = ctx.sweep_autotune(5)
    cands = t.candidates()
    print(f"  autotune: chose {t.threads_per_block} threads, schedule {t.schedule}, loads {t.loads}; {cands}",
          flush=True)
    assert t.n_candidates == 32
    assert {"threads": t.threads_per_block, "schedule": t.schedule, "loads": t.loads} in \
        [{k: c[k] for k in ("threads", "schedule", "loads")} for c in cands]
    U1, V1, _ = _solve(ctx, B, precision=abi.UFE_MIXED)
    for b in range(len(wls)):
        assert np.array_equal(U0[b], U1[b]) and np.array_equal(V0[b], V1[b]), b
    # every other shape gives the same bits too
    for shape in ((64, 1, 2), (512, 0, 1), (256, 0, 3)):
        ctx.sweep_tuning_set(*shape)
        U2, V2, _ = _solve(ctx, B, precision=abi.UFE_MIXED)
        for b in range(len(wls)):
            assert np.array_equal(U0[b], U2[b]) and np.array_equal(V0[b], V2[b]), (shape, b)
    ctx.sweep_tuning_set(t.threads_per_block, t.schedule, t.loads)


def test_batch_multigrid_matches_fp64(ens):
    """The union of the members aggregated as one hierarchy: every member
    agrees with the FP64 batch, and the batch's graph replays (the coarse
    levels recorded inline) equal eager launches (the coarse levels replayed
    from the multigrid's own graph) bit for bit."""
    ctx, B, wls = ens
    U0, V0, _ = _solve(ctx, B, precision=abi.UFE_FP64)
    kw = dict(precision=abi.UFE_MIXED, precond_type=abi.UFE_PRECOND_MULTIGRID, precond_sweeps=2, gmres_restart=20)
    Ug, Vg, sg = _solve(ctx, B, use_graph=True, **kw)
    Ue, Ve, se = _solve(ctx, B, use_graph=False, **kw)
    print(f"  multigrid batch: cycles {sg.cycles} sweeps {sg.sweeps} {sg.seconds:.3f}s; "
          f"hierarchy {ctx.multigrid_info()['points']}", flush=True)
    assert sg.cycles == se.cycles
    for b in range(len(wls)):
        e = max(_rel(Ug[b], U0[b]), _rel(Vg[b], V0[b]))
        print(f"  member {b}: multigrid vs fp64 rel L2 {e:.2e}", flush=True)
        assert e <= 1e-7, (b, e)
        assert np.array_equal(Ug[b], Ue[b]) and np.array_equal(Vg[b], Ve[b]), b
    with pytest.raises(abi.UFEError):
        ctx.batch_linear_solve(precision=abi.UFE_FP64, precond_type=abi.UFE_PRECOND_MULTIGRID)
