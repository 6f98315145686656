"""ice_physical_properties on the GPU (src/general_ice_model_data_module.f90:298-462)
against the oracle: the Arrhenius flow factor, pressure-melting point, heat
capacity and conductivity on Aa, Ti_Ac and the Ac flow factors, the fixed
benchmark values, and a whole dynamics step whose A comes from Ti."""
import numpy as np
import pytest

from ufemism_amd import abi, geometry
from ufemism_amd.dynamics import Step

pytestmark = pytest.mark.gpu

FIELDS = ("A_flow", "A_flow_mean", "Ti_pmp", "Cpi", "Ki", "Ti_Ac", "A_flow_Ac", "A_flow_mean_Ac")
TOL_ELEM = 1e-14   # elementwise relative; exp() of the device and of glibc may differ in the last bit


def _oracle():
    from oracle import pyoracle
    return pyoracle


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="module")
def marine():
    print("[phys] marine20k with an englacial temperature field", flush=True)
    wl = geometry.marine(20000, seed=5, picard=8)
    wl.params.physical_properties = 1
    ref = _oracle().run_step(wl)
    return wl, ref


def test_properties_match_oracle(ctx, marine):
    wl, ref = marine
    om = ref["mesh"]
    po = _oracle()
    ice = po.OracleIce(om)
    ice.set(Hi=wl.Hi, Hb=wl.Hb, SL=wl.SL, Ti=wl.temperature())
    ice.update_general_ice_model_data()
    ice.ice_physical_properties(wl.params, 1)
    Step(ctx, wl)
    ctx.update_general_ice_model_data()
    assert np.array_equal(ctx.get("mask_sheet"), ice["mask_sheet"])
    assert np.array_equal(ctx.get("mask_sheet_Ac"), ice["mask_sheet_Ac"])
    for k in FIELDS:
        g, o = ctx.get(k), ice[k]
        e = float(np.max(np.abs(g - o) / np.maximum(np.abs(o), 1e-300)))
        print(f"  {k}: bit-exact={np.array_equal(g, o)} max elementwise rel {e:.1e}", flush=True)
        assert e <= TOL_ELEM, (k, e)
    # both branches of the law are exercised
    assert ice["Ti"].min() < 263.15 < ice["Ti"].max()


def test_standalone_entry_equals_update_path(ctx, marine):
    wl, _ = marine
    Step(ctx, wl)
    ctx.update_general_ice_model_data()
    a = {k: ctx.get(k) for k in FIELDS}
    for k in ("A_flow", "A_flow_Ac", "A_flow_mean", "A_flow_mean_Ac"):
        ctx.fill(k, 0.0)
    ctx.ice_physical_properties(1)
    for k in FIELDS:
        assert np.array_equal(ctx.get(k), a[k]), k


def test_benchmark_fixed_values_and_bad_modes(ctx, marine):
    wl, _ = marine
    Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.ice_physical_properties(2)
    for k in ("A_flow", "A_flow_Ac", "A_flow_mean", "A_flow_mean_Ac"):
        assert np.all(ctx.get(k) == 1e-16), k
    assert np.all(ctx.get("Cpi") == 2009.0)
    assert np.all(ctx.get("Ki") == 2.1 * 31556943.36)
    po = _oracle()
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    ice.set(Hi=wl.Hi, Hb=wl.Hb, SL=wl.SL)
    ice.update_general_ice_model_data()
    ice.ice_physical_properties(wl.params, 2)
    assert np.array_equal(ctx.get("Ti_pmp"), ice["Ti_pmp"])
    with pytest.raises(abi.UFEError):
        ctx.ice_physical_properties(0)
    with pytest.raises(abi.UFEError):
        ctx.ice_physical_properties(3)


def test_dynamics_step_with_arrhenius_flow_factor_matches_oracle(ctx, marine):
    wl, ref = marine
    out = Step(ctx, wl).run()
    print(f"  sweeps gpu={out['sweeps']} oracle={ref['sweeps']}", flush=True)
    assert out["sweeps"] == ref["sweeps"]
    for k in ("U_SIA", "D_SIA_Ac", "U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Hi", "dHi_dt"):
        e = _rel(out[k], ref[k])
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-9, (k, e)
    # the temperature field changes the answer: A is not the workload's constant
    A = ctx.get("A_flow_mean")
    assert A.min() < 0.5 * A.max()
