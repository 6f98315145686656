"""The step after the path on the GPU against the oracle: solve_SIA_3D, the
critical time steps (a min-reduction: bit-exact on identical inputs), the
upwind derivatives and a multi-step run of the dynamics time loop with the
reference's time-step selection."""
import numpy as np
import pytest

from ufemism_amd import geometry
from ufemism_amd.dynamics import RegionClock, Step, TimeConfig, initial_fields, run_dynamics

pytestmark = pytest.mark.gpu


def _oracle():
    from oracle import pyoracle
    return pyoracle


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def _oracle_state(wl):
    po = _oracle()
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    for k, v in initial_fields(wl, om["VAaAc"], wl.params.nZ).items():
        if k in ice.arr:
            ice[k][...] = v
    ice.update_general_ice_model_data()
    ice.solve_SIA(wl.params)
    ice.solve_SSA(wl.params)
    ice.solve_SIA_3D(wl.params)
    return om, ice


@pytest.mark.parametrize("name", ["dome8k", "marine20k"])
def test_sia3d_and_critical_steps_match_oracle(ctx, name):
    print(f"[sia3d] {name}", flush=True)
    wl = {"dome8k": lambda: geometry.dome(8000, seed=1, picard=6),
          "marine20k": lambda: geometry.marine(20000, seed=5, picard=6)}[name]()
    om, ice = _oracle_state(wl)
    Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    ctx.solve_SSA()
    ctx.solve_SIA_3D()
    for k in ("U_3D", "V_3D", "W_3D"):
        e = _rel(ctx.get(k), ice[k])
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-10, (k, e)
    g = ctx.critical_timesteps()
    o = ice.critical_timesteps(wl.params)
    print(f"  critical steps gpu {g} oracle {o}", flush=True)
    assert np.allclose(g, o, rtol=1e-10)
    # identical inputs: the min-reduction gives the oracle's numbers exactly
    for k in ("D_SIA_Ac", "U_SSA", "V_SSA", "U_3D", "V_3D"):
        ctx.set(k, ice[k])
    assert ctx.critical_timesteps() == o


def test_upwind_derivatives_match_oracle(ctx):
    import torch
    po = _oracle()
    wl = geometry.marine(20000, seed=5, picard=2)
    Step(ctx, wl)
    om = po.OracleMesh(wl.mesh).build()
    nV, nz = wl.mesh.nV, 4
    rng = np.random.default_rng(11)
    U = rng.standard_normal((nV, nz)) * 100
    V = rng.standard_normal((nV, nz)) * 100
    U[::13, 2] = V[::13, 2] = 0.0
    d = rng.standard_normal((nV, nz)) * 10
    ox, oy, omiss, _, _ = po.upwind_derivatives_3D(om, om.geometry["R"], U, V, d)
    dev = [torch.from_numpy(np.asarray(a).ravel(order="F").copy()).to("cuda") for a in (U, V, d)]
    outx = torch.zeros(nV * nz, dtype=torch.float64, device="cuda")
    outy = torch.zeros_like(outx)
    torch.cuda.synchronize()
    miss = ctx.upwind_derivatives_3D(*(t.data_ptr() for t in dev), outx.data_ptr(), outy.data_ptr(), nz)
    gx = outx.cpu().numpy().reshape((nV, nz), order="F")
    gy = outy.cpu().numpy().reshape((nV, nz), order="F")
    print(f"  upwind: missing gpu {miss} oracle {omiss}; bit-exact {np.array_equal(gx, ox) and np.array_equal(gy, oy)}",
          flush=True)
    assert miss == omiss
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)


def test_time_loop_matches_oracle(ctx):
    """Twenty model years of the dynamics time loop on a 2k dome: the same
    sequence of actions, steps within rounding, the same final thickness."""
    po = _oracle()
    wl = geometry.dome(2000, seed=3, picard=5)
    cfg = TimeConfig(dt_thermo=5.0)
    ref_log, ice, _ = po.run_dynamics(wl, 20.0, RegionClock(cfg))
    Step(ctx, wl)
    log = run_dynamics(ctx, 20.0, RegionClock(cfg))
    print(f"  steps gpu {len(log)} oracle {len(ref_log)}", flush=True)
    assert len(log) == len(ref_log)
    for (t, dt, do, st), (rt, rdt, rdo, rst) in zip(log, ref_log):
        assert do == rdo
        assert abs(t - rt) <= 1e-9 * max(1.0, abs(rt)) and abs(dt - rdt) <= 1e-9 * max(1.0, abs(rdt))
        if st is not None:
            assert list(st.sweeps)[: st.picard_iterations] == list(rst.sweeps)[: rst.picard_iterations]
    for k in ("Hi", "U_SSA", "V_SSA", "U_3D", "W_3D"):
        e = _rel(ctx.get(k), ice[k])
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-9, (k, e)
    assert any(do["thermo"] for _, _, do, _ in log)
