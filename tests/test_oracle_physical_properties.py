"""The oracle's ice_physical_properties (src/general_ice_model_data_module.f90:298-462),
pinned by published ice physics and by closed forms rather than by its own
formulas: the two Arrhenius branches meet at 263.15 K, their activation
energies come back out of log-ratios, the flow factor, heat capacity and
conductivity sit at the published values near -10 C, the trapezoid mean is
exact on isothermal and linear-in-zeta profiles, Ti_Ac of a linear field is
the field at the edge midpoint, and the shelf branch equals the sheet law at
(Ti_surface + 271.15 K) / 2."""
import numpy as np
import pytest

from ufemism_amd import geometry
from ufemism_amd.config import DynParams
from ufemism_amd.fields import ZETA_DEFAULT

po = pytest.importorskip("oracle.pyoracle")

SEC_PER_YEAR = 31556943.36
R_GAS = 8.314


@pytest.fixture(scope="module")
def state():
    wl = geometry.marine(3000, seed=5, picard=2)
    wl.params = DynParams(physical_properties=1)
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    ice.set(Hi=wl.Hi, Hb=wl.Hb, SL=wl.SL)
    ice.update_general_ice_model_data()
    return wl, om, ice


def _run(ice, p, Ti, mode=1):
    ice["Ti"][...] = Ti
    ice.ice_physical_properties(p, mode)
    return {k: ice[k].copy() for k in ("A_flow", "A_flow_mean", "Ti_pmp", "Cpi", "Ki",
                                       "Ti_Ac", "A_flow_Ac", "A_flow_mean_Ac")}


def _uniform(ice, p, T):
    nV = ice["Hi"].shape[0]
    return _run(ice, p, np.full((nV, p.nZ), T))


def test_arrhenius_branches_meet_and_give_their_activation_energies(state):
    wl, om, ice = state
    p = wl.params
    A = {T: _uniform(ice, p, T)["A_flow"][0, 0] for T in (240.0, 250.0, 263.15 - 1e-9, 263.15, 268.0, 272.0)}
    # the two branches join within 3 % at the 263.15 K switch
    lo, hi = A[263.15 - 1e-9], A[263.15]
    assert abs(hi / lo - 1.0) < 0.03, (lo, hi)
    # d ln A / d(1/T) = -Q/R in each branch: 60 kJ/mol below, 139 kJ/mol above
    Q_low = -R_GAS * np.log(A[250.0] / A[240.0]) / (1 / 250.0 - 1 / 240.0)
    Q_high = -R_GAS * np.log(A[272.0] / A[268.0]) / (1 / 272.0 - 1 / 268.0)
    assert abs(Q_low - 6.0e4) < 1e-6 * 6.0e4 and abs(Q_high - 13.9e4) < 1e-6 * 13.9e4
    # A(-10 C) ~ 3.5e-25 Pa^-3 s^-1 (Cuffey and Paterson 2010) = 1.1e-17 Pa^-3 yr^-1
    assert 0.5e-17 < hi < 2.5e-17
    # warmer ice is softer within each branch; the law steps down by 0.2 % at the switch
    for Ts in ((240.0, 250.0, 263.15 - 1e-9), (263.15, 268.0, 272.0)):
        assert all(A[a] < A[b] for a, b in zip(Ts, Ts[1:]))
    assert hi < lo


def test_heat_capacity_conductivity_and_melting_point(state):
    wl, om, ice = state
    p = wl.params
    out = _uniform(ice, p, 263.15)
    # specific heat of ice at -10 C ~ 2030 J kg^-1 K^-1 (within 1 %)
    assert abs(out["Cpi"][0, 0] / 2030.0 - 1) < 0.01
    # Ritz (1987): k = 9.828 exp(-0.0057 T) W m^-1 K^-1, here per year with the
    # prefactor rounded to 3.101e8 (0.013 % low)
    for T in (230.0, 263.15, 273.0):
        Ki = _uniform(ice, p, T)["Ki"][0, 0] / SEC_PER_YEAR
        assert abs(Ki / (9.828 * np.exp(-0.0057 * T)) - 1) < 2e-4
    assert abs(out["Ki"][0, 0] / SEC_PER_YEAR / 2.3 - 1) < 0.1   # ~2.3 W/m/K at -10 C
    # pressure melting point: T0 at the surface, 0.87 K lower per km of ice at the base
    H = ice["Hi"]
    assert np.array_equal(out["Ti_pmp"][:, 0], np.full_like(H, 273.16))
    thick = H > 500
    drop = (273.16 - out["Ti_pmp"][thick, -1]) / (H[thick] / 1000.0)
    assert np.allclose(drop, 0.87, rtol=1e-12)
    # linear in zeta
    z = np.asarray(p.zeta)
    assert np.allclose(out["Ti_pmp"], 273.16 - 0.87e-3 * H[:, None] * z[None, :], rtol=0, atol=1e-9)


def test_vertical_means_and_shelf_branch(state):
    wl, om, ice = state
    p = wl.params
    nV = ice["Hi"].shape[0]
    sheet = ice["mask_sheet"] == 1
    assert sheet.any() and (~sheet).any()
    # isothermal columns: the trapezoid mean of a constant is that constant
    out = _uniform(ice, p, 255.0)
    A255 = out["A_flow"][0, 0]
    assert np.allclose(out["A_flow_mean"][sheet], A255, rtol=1e-14)
    # shelf points use the law at (Ti(surface) + 271.15) / 2; with Ti = 271.15 that is A(271.15)
    out = _uniform(ice, p, 271.15)
    assert np.allclose(out["A_flow_mean"][~sheet], out["A_flow"][0, 0], rtol=1e-15)
    # a general column: the mean is numpy's trapezoid of the A profile over zeta
    Ti = wl.temperature()
    out = _run(ice, p, Ti)
    z = np.asarray(p.zeta)
    trap = np.trapezoid(out["A_flow"][sheet], z, axis=1)
    assert np.allclose(out["A_flow_mean"][sheet], trap, rtol=1e-13)
    # shelf column with a 251.15 K surface: (251.15 + 271.15) / 2 = 261.15 K, the low branch
    T2 = np.full((nV, p.nZ), 261.15)
    T2[:, 0] = 251.15
    out2 = _run(ice, p, T2)
    assert np.allclose(out2["A_flow_mean"][~sheet], out2["A_flow"][0, 1], rtol=1e-15)


def test_ac_temperature_and_flow_factor(state):
    wl, om, ice = state
    p = wl.params
    V, VAc = om.mesh.V, om["VAc"]
    # a field linear in x and depth: its Ac value is the field at the edge midpoint
    z = np.asarray(p.zeta)
    Ti = 250.0 + 1e-5 * V[:, 0][:, None] + 5.0 * z[None, :]
    out = _run(ice, p, Ti)
    exact = 250.0 + 1e-5 * VAc[:, 0][:, None] + 5.0 * z[None, :]
    assert np.allclose(out["Ti_Ac"], exact, rtol=0, atol=1e-9)
    # the Ac flow factor follows the same law as the Aa one
    out = _uniform(ice, p, 266.0)
    assert np.all(out["A_flow_Ac"] == out["A_flow"][0, 0])
    sheetA = ice["mask_sheet_Ac"] == 1
    assert sheetA.any() and (~sheetA).any()
    assert np.allclose(out["A_flow_mean_Ac"][sheetA], out["A_flow"][0, 0], rtol=1e-14)
    out = _uniform(ice, p, 271.15)
    assert np.allclose(out["A_flow_mean_Ac"][~sheetA], out["A_flow"][0, 0], rtol=1e-15)


def test_benchmark_fixed_values(state):
    wl, om, ice = state
    p = wl.params
    nV = ice["Hi"].shape[0]
    ice["Ti"][...] = 250.0
    ice.ice_physical_properties(p, 2)
    # EISMINT / Halfar / Bueler values (src/general_ice_model_data_module.f90:333-343)
    assert np.all(ice["A_flow"] == 1e-16) and np.all(ice["A_flow_Ac"] == 1e-16)
    assert np.all(ice["A_flow_mean"] == 1e-16) and np.all(ice["A_flow_mean_Ac"] == 1e-16)
    assert np.allclose(ice["Ki"] / SEC_PER_YEAR, 2.1, rtol=1e-15)
    assert np.all(ice["Cpi"] == 2009.0)
    z = np.asarray(p.zeta)
    assert np.allclose(ice["Ti_pmp"], 273.16 - 0.87e-3 * ice["Hi"][:, None] * z[None, :], rtol=0, atol=1e-9)
    assert ice["A_flow"].shape == (nV, len(ZETA_DEFAULT))


def test_workload_temperature_spans_both_branches(state):
    wl, om, ice = state
    Ti = wl.temperature()
    assert Ti.min() < 263.15 < Ti.max()
    pmp = 273.16 - 8.7e-4 * wl.Hi[:, None] * np.asarray(wl.params.zeta)[None, :]
    assert np.all(Ti < pmp)
