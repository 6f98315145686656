"""The CPU oracle's physics, pinned by closed forms, a direct sparse solve,
conservation laws and the reference's failure paths."""
import ctypes as ct

import numpy as np
import pytest

from ufemism_amd import geometry, meshgen
from ufemism_amd.config import DynParams, tau_c_from_beta
from ufemism_amd.fields import ZETA_DEFAULT

po = pytest.importorskip("oracle.pyoracle")

RHO_G = 910.0 * 9.81


def _vert(p, f, name):
    lib = po.load()
    ps = po.params_struct(p)
    f = np.ascontiguousarray(f, float)
    if name == "average":
        return lib.om_vertical_average(ct.byref(ps), f.ctypes.data_as(ct.c_void_p))
    out = np.zeros_like(f)
    lib.om_vertical_integrate(ct.byref(ps), f.ctypes.data_as(ct.c_void_p), out.ctypes.data_as(ct.c_void_p))
    return out


def test_vertical_integrals_trapezoid():
    p = DynParams()
    z = ZETA_DEFAULT
    f = 2.0 + 3.0 * z                               # trapezoid is exact on linear f
    assert np.isclose(_vert(p, f, "average"), 3.5, rtol=1e-15)
    I = _vert(p, f, "integrate")
    exact = -((2.0 * (1 - z)) + 1.5 * (1 - z ** 2))  # -int_z^1 f
    assert np.allclose(I, exact, rtol=1e-14, atol=1e-15)
    # I_zeta = average(integrate(zeta^3)) tends to -1/(n+2) = -0.2 on fine grids
    fine = DynParams(zeta=np.linspace(0, 1, 32))
    Iz = _vert(fine, _vert(fine, fine.zeta ** 3, "integrate"), "average")
    assert abs(Iz + 0.2) < 1e-3


@pytest.fixture(scope="module")
def dome_state():
    wl = geometry.dome(3000, seed=21, picard=3)
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    ice.set(Hi=wl.Hi, Hb=wl.Hb, SL=wl.SL, A_flow=wl.A_flow, A_flow_Ac=wl.A_flow,
            A_flow_mean=wl.A_flow, A_flow_mean_Ac=wl.A_flow)
    ice["tau_c_override_AaAc"][:] = tau_c_from_beta(wl.beta)
    ice.update_general_ice_model_data()
    return wl, om, ice


def test_sia_closed_form_on_the_dome(dome_state):
    """With uniform A and m_enh = 1 the SIA diffusivity separates:
    D3D(zeta) = 2 A (rho g)^n H^(n+1) |grad s|^(n-1) J(zeta), J = -int_zeta^1 z^n dz,
    cut off at -1e5, and D_SIA = H * (trapezoid mean of D3D)."""
    wl, om, ice = dome_state
    ice.solve_SIA(wl.params)
    z = ZETA_DEFAULT
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (z[1:] ** 3 + z[:-1] ** 3) * np.diff(z))])
    J = -(cum[-1] - cum)
    H = ice["Hi_Ac"]
    g2 = ice["dHs_dp_Ac"] ** 2 + ice["dHs_do_Ac"] ** 2
    D3 = 2 * wl.A_flow * RHO_G ** 3 * H[:, None] ** 4 * g2[:, None] * J[None, :]
    D3 = np.maximum(D3, -1e5)
    D = H * np.sum(0.5 * (D3[:, 1:] + D3[:, :-1]) * np.diff(z)[None, :], axis=1)
    sheet = ice["mask_sheet_Ac"] == 1
    assert sheet.sum() > 100
    assert np.allclose(ice["D_SIA_Ac"][sheet], D[sheet], rtol=1e-10, atol=1e-12)
    assert np.all(ice["D_SIA_Ac"][~sheet] == 0.0)
    # closed form without the cut-off, where it is inactive: I_zeta of the grid
    Iz = np.sum(0.5 * (J[1:] + J[:-1]) * np.diff(z))
    free = sheet & (D3.min(axis=1) > -1e5)
    assert np.allclose(ice["D_SIA_Ac"][free],
                       2 * wl.A_flow * RHO_G ** 3 * H[free] ** 5 * g2[free] * Iz, rtol=1e-10)
    # flux runs downhill: U_p = D2D * ds/dp with D2D <= 0
    assert np.all(ice["Up_SIA_Ac"][sheet] * ice["dHs_dp_Ac"][sheet] <= 0)


def test_yield_stress_closed_forms(dome_state):
    wl, om, ice = dome_state
    lib = po.load()
    Hb = ice["Hb"]
    for hb, phi in ((0.0, 20.0), (-1000.0, 5.0), (-500.0, 12.5), (1000.0, 20.0)):
        ice["Hb"][:] = hb
        ice["Hb_Ac"][:] = hb
        lib.om_basal_yield_stress(ct.byref(om.s), ct.byref(ice.s))
        H = np.maximum(0.1, ice["Hi"])
        lam = min(1.0, max(0.0, 1.0 - hb / 1000.0))
        tau = np.tan(np.pi / 180 * phi) * (RHO_G * H - 0.96 * RHO_G * H * lam)
        n = om.mesh.nV
        assert np.allclose(ice["phi_fric_AaAc"][:n], phi, rtol=1e-14)
        assert np.allclose(ice["tau_c_AaAc"][:n], tau, rtol=1e-12)
    ice["Hb"][:] = Hb
    ice["Hb_Ac"][:] = 0.0


def test_viscosity_closed_form_for_linear_flow(dome_state):
    wl, om, ice = dome_state
    P = om["VAaAc"]
    a, b, c, d = 2e-3, -1e-3, 5e-4, -3e-3
    ice["U_SSA_AaAc"][:] = a * P[:, 0] + b * P[:, 1]
    ice["V_SSA_AaAc"][:] = c * P[:, 0] + d * P[:, 1]
    ice["A_flow_mean_AaAc"][:] = 1e-16
    ice["Hi_AaAc"][:] = 500.0
    p = DynParams(m_enh_ssa=0.7)
    po.load().om_SSA_effective_viscosity(ct.byref(om.s), ct.byref(ice.s), ct.byref(po.params_struct(p)))
    eps2 = a * a + d * d + a * d + 0.25 * (b + c) ** 2 + 1e-12
    eta = (0.7 * 0.5 * 1e-16) ** (-1 / 3) * eps2 ** (-1 / 3)
    ok = np.isfinite(ice["eta_AaAc"])
    assert np.allclose(ice["eta_AaAc"][ok], eta, rtol=1e-8)
    assert np.allclose(ice["N_AaAc"][ok], eta * 500.0, rtol=1e-8)


def test_sliding_closed_form(dome_state):
    wl, om, ice = dome_state
    ice["tau_c_AaAc"][:] = 7.0e4
    ice["U_SSA_AaAc"][:] = 0.0
    ice["V_SSA_AaAc"][:] = 0.0
    po.load().om_SSA_sliding_term(ct.byref(om.s), ct.byref(ice.s))
    # (1e-6)^(-0.35) / 100^0.3 = 10^2.1 / 10^0.6 = 10^1.5
    assert np.allclose(ice["S_AaAc"], 7.0e4 * 10 ** 1.5, rtol=1e-12)
    ice["U_SSA_AaAc"][:] = 300.0
    ice["V_SSA_AaAc"][:] = 400.0
    po.load().om_SSA_sliding_term(ct.byref(om.s), ct.byref(ice.s))
    assert np.allclose(ice["S_AaAc"], 7.0e4 * (1e-6 + 500.0 ** 2) ** (-0.35) / 100 ** 0.3, rtol=1e-12)


def _small_ssa():
    wl = geometry.dome(600, seed=5, picard=1)
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    ice.set(Hi=wl.Hi, Hb=wl.Hb, SL=wl.SL, A_flow=wl.A_flow, A_flow_Ac=wl.A_flow,
            A_flow_mean=wl.A_flow, A_flow_mean_Ac=wl.A_flow)
    ice["tau_c_override_AaAc"][:] = tau_c_from_beta(wl.beta)
    ice.update_general_ice_model_data()
    return wl, om, ice


def _assembled_ssa(om, ice, fixed=None):
    """The linearised system of the last Picard iteration, assembled: the SSA
    stencil on SOR points, Neumann rows on edge points (edge first, as the
    reference tests it first), and identity rows holding the current values on
    the points in `fixed` (not on edges)."""
    import scipy.sparse as sp
    n = om.nVAaAc
    nc, C = om["nCAaAc"], om["CAaAc"] - 1
    edge = np.concatenate([om.mesh.edge_index > 0, om["edge_index_Ac"] > 0])
    fixed = np.zeros(n, bool) if fixed is None else fixed
    I, J, X = [], [], []
    bu, bv = np.zeros(n), np.zeros(n)
    for i in range(n):
        k = nc[i]
        if i < 4 or edge[i]:
            nb = [j for j in C[i, :k] if (i < 4 or not edge[j])]
            for comp in (0, 1):
                I.append(comp * n + i); J.append(comp * n + i); X.append(1.0)
                for j in nb:
                    I.append(comp * n + i); J.append(comp * n + j); X.append(-1.0 / len(nb))
            continue
        if fixed[i]:
            I += [i, n + i]; J += [i, n + i]; X += [1.0, 1.0]
            bu[i], bv[i] = ice["U_SSA_AaAc"][i], ice["V_SSA_AaAc"][i]
            continue
        sxy = om["Nxy_AaAc"][i, k] + om["Nxy_AaAc"][i, :k].sum()
        for c in range(k):
            j = C[i, c]
            I += [i, n + i]; J += [j, n + j]
            X += [4 * om["Nxx_AaAc"][i, c] + om["Nyy_AaAc"][i, c], 4 * om["Nyy_AaAc"][i, c] + om["Nxx_AaAc"][i, c]]
        I += [i, i, n + i, n + i]; J += [i, n + i, n + i, i]
        X += [ice["eu_i_AaAc"][i], 3 * sxy, ice["ev_i_AaAc"][i], 3 * sxy]
        bu[i], bv[i] = ice["RHSx_AaAc"][i], ice["RHSy_AaAc"][i]
    return sp.csr_matrix((X, (I, J)), shape=(2 * n, 2 * n)), np.concatenate([bu, bv])


def _direct_error(om, ice, fixed=None):
    import scipy.sparse.linalg as spl
    A, b = _assembled_ssa(om, ice, fixed)
    x = spl.spsolve(A.tocsc(), b)
    n = om.nVAaAc
    u, v = ice["U_SSA_AaAc"], ice["V_SSA_AaAc"]
    return np.sqrt(np.sum((x[:n] - u) ** 2 + (x[n:] - v) ** 2)) / np.sqrt(np.sum(x ** 2))


def test_sor_fixed_point_is_the_direct_solution():
    """SOR run to a tiny tolerance reaches the solution of the assembled
    linear system (SSA stencil on SOR points, Neumann rows on edge points)."""
    wl, om, ice = _small_ssa()
    p = wl.params
    p.SSA_max_outer_loops, p.SSA_max_residual_UV, p.SSA_max_inner_loops = 1, 1e-11, 200000
    st = ice.solve_SSA(p)
    assert st.status == 0 and st.sweeps[0] < 200000
    err = _direct_error(om, ice)
    assert err < 1e-8, err


def _linearised(om, ice, p):
    """One call of the oracle's solve_SSA_linearised on the fields already set."""
    ps = po.params_struct(p)
    did, lm, ns = ct.c_int(0), ct.c_double(0), ct.c_int(0)
    r = po.load().om_solve_SSA_linearised(ct.byref(om.s), ct.byref(ice.s), ct.byref(ps), ct.byref(did),
                                          ct.byref(lm), ct.byref(ns))
    return r, did.value, lm.value, ns.value


def _bare_state(n=1000, seed=5):
    m = meshgen.generate_mesh(geometry.dome_resolution(), n, seed=seed)
    om = po.OracleMesh(m).build()
    assert om.colour_status == 0
    edge = np.concatenate([m.edge_index > 0, om["edge_index_Ac"] > 0])
    return om, po.OracleIce(om), edge


@pytest.mark.parametrize("H0", [1500.0, 0.05])
def test_linearised_ssa_plug_flow_closed_form(H0):
    """Uniform grounded ice with uniform eta, S and surface slope: the second-
    derivative weights annihilate constants and Neumann rows keep them, so the
    linearised SSA (src/ice_dynamics_module.f90:581-595, 638-644)
        4u_xx + u_yy + 3v_xy - S u / (max(0.1,H) eta) = rho g ds/dx / eta
    has the exact plug solution u = -rho g max(0.1,H) ds/dx / S and likewise
    v with ds/dy.  Slopes of different size in x and y catch a u/v swap; the
    thin-ice case catches a missing max(0.1, H); rho g is 910 * 9.81."""
    om, ice, edge = _bare_state()
    n = om.nVAaAc
    gx, gy, S, eta = -2.0e-3, 7.0e-4, 5.0e3, 1.0e6
    ice["Hi_AaAc"][:] = H0
    ice["Hb_AaAc"][:] = 0.0                 # SL = Hb = 0: grounded at any thickness
    ice["SL_AaAc"][:] = 0.0
    ice["eta_AaAc"][:] = eta
    ice["S_AaAc"][:] = S
    ice["dHs_dx_shelf_AaAc"][:] = gx
    ice["dHs_dy_shelf_AaAc"][:] = gy
    ice["U_SSA_AaAc"][:] = 0.0
    ice["V_SSA_AaAc"][:] = 0.0
    u0 = -RHO_G * max(0.1, H0) * gx / S
    v0 = -RHO_G * max(0.1, H0) * gy / S
    p = DynParams(SSA_SOR_omega=1.3, SSA_max_inner_loops=100000, SSA_max_residual_UV=1e-12 * abs(u0))
    r, did, _, sweeps = _linearised(om, ice, p)
    assert r == 0 and did == 0 and sweeps < 100000
    assert np.allclose(ice["U_SSA_AaAc"], u0, rtol=1e-9, atol=0)
    assert np.allclose(ice["V_SSA_AaAc"], v0, rtol=1e-9, atol=0)
    assert n == ice["U_SSA_AaAc"].size


def _lhs_probe(om, ice, u, v):
    """The oracle's left-hand sides at fixed (u, v): one sweep with omega = 0
    updates nothing on SOR points, so LHSx/LHSy hold the discrete operator."""
    ice["U_SSA_AaAc"][:] = u
    ice["V_SSA_AaAc"][:] = v
    _linearised(om, ice, DynParams(SSA_SOR_omega=0.0, SSA_max_inner_loops=1, SSA_max_residual_UV=0.0))
    return ice["LHSx_AaAc"].copy(), ice["LHSy_AaAc"].copy()


def _mixed_ice(om, seed=3):
    """Non-uniform thickness, viscosity and friction; the points with x < 0
    float (Hb = -1000 m, SL = 0, H = 500 m < 1000 * 1028/910 m), the others are
    grounded on a bed at sea level."""
    rng = np.random.default_rng(seed)
    n = om.nVAaAc
    x = om["VAaAc"][:, 0]
    ice = po.OracleIce(om)
    flo = x < 0
    ice["Hi_AaAc"][:] = np.where(flo, 500.0, 1000.0 + 500.0 * rng.random(n))
    ice["Hb_AaAc"][:] = np.where(flo, -1000.0, 0.0)
    ice["SL_AaAc"][:] = 0.0
    ice["eta_AaAc"][:] = 1e8 * (1.0 + rng.random(n))
    ice["S_AaAc"][:] = 1e4 * (1.0 + rng.random(n))
    ice["dHs_dx_shelf_AaAc"][:] = 1e-3 * rng.standard_normal(n)
    ice["dHs_dy_shelf_AaAc"][:] = 1e-3 * rng.standard_normal(n)
    return ice, flo


@pytest.mark.parametrize("n", [1000, 4000])
def test_linearised_ssa_operator_on_manufactured_quadratics(n):
    """The discrete operator against the PDE on u = (x/L)^2, v = (y/L)^2:
    4u_xx + u_yy = 8/L^2 and 4v_yy + v_xx = 8/L^2 (the swapped weights
    4Nyy+Nxx on u would give 2/L^2), minus the friction term
    S u / (max(0.1,H) eta) on grounded points only (floating points carry
    none).  The second-derivative weights are consistent to first order, so
    the median error over interior points is bounded and falls with the
    resolution.  The right-hand sides are rho g ds/dx / eta pointwise."""
    L = 800e3
    errs = []
    for nn in (n, 4 * n):
        om, _, edge = _bare_state(nn, seed=7)
        ice, flo = _mixed_ice(om)
        X = om["VAaAc"]
        u, v = (X[:, 0] / L) ** 2, (X[:, 1] / L) ** 2
        lx, ly = _lhs_probe(om, ice, u, v)
        beta = np.where(flo, 0.0, ice["S_AaAc"] / (np.maximum(0.1, ice["Hi_AaAc"]) * ice["eta_AaAc"]))
        inner = ~edge
        ex = np.median(np.abs(lx - (8 / L ** 2 - beta * u))[inner]) * L ** 2 / 8
        ey = np.median(np.abs(ly - (8 / L ** 2 - beta * v))[inner]) * L ** 2 / 8
        errs.append(max(ex, ey))
        assert max(ex, ey) < 0.02, (nn, ex, ey)
        # the friction term is of the same size as the curvature term here, so a
        # friction term on floating points (or none on grounded ones) is seen
        assert np.median((beta * u)[inner & ~flo]) * L ** 2 / 8 > 0.05
        assert np.allclose(ice["RHSx_AaAc"], RHO_G * ice["dHs_dx_shelf_AaAc"] / ice["eta_AaAc"], rtol=1e-14)
        assert np.allclose(ice["RHSy_AaAc"], RHO_G * ice["dHs_dy_shelf_AaAc"] / ice["eta_AaAc"], rtol=1e-14)
    assert errs[1] < 0.5 * errs[0], errs


def test_linearised_ssa_cross_term_is_the_reference_curvature_bug():
    """u = 0, v = xy/L^2: the PDE's 3 v_xy = 3/L^2, but the reference takes the
    cross curvature from get_mesh_curvatures_vertex_AaAc, which multiplies
    every weight by the home value (src/mesh_ArakawaC_module.f90:647-656), so
    the term is v_i times the weights' sum, which vanishes to rounding; the
    coupling lands in the u-equation (LHSx), and LHSy sees none of it."""
    L = 800e3
    om, _, edge = _bare_state(2000, seed=8)
    ice, flo = _mixed_ice(om)
    X = om["VAaAc"]
    v = X[:, 0] * X[:, 1] / L ** 2
    lx, ly = _lhs_probe(om, ice, np.zeros(om.nVAaAc), v)
    inner = ~edge
    assert np.max(np.abs(lx[inner])) < 1e-6 * 3 / L ** 2
    # the same v shifted by a constant changes LHSx only through that
    # rounding-size weight sum, never by O(3/L^2)
    lx2, _ = _lhs_probe(om, ice, np.zeros(om.nVAaAc), v + 1.0)
    assert np.max(np.abs(lx2[inner] - lx[inner])) < 1e-6 * 3 / L ** 2
    # v-equation: 4v_yy + v_xx = 0 for v = xy, so LHSy is the friction term alone
    beta = np.where(flo, 0.0, ice["S_AaAc"] / (np.maximum(0.1, ice["Hi_AaAc"]) * ice["eta_AaAc"]))
    scale = np.median((beta * np.abs(v))[inner & ~flo])
    assert np.median(np.abs(ly + beta * v)[inner]) < 0.05 * scale


def _rn_state():
    wl, om, ice = _small_ssa()
    x, y = om.mesh.V[:, 0], om.mesh.V[:, 1]
    A = om["Aci"] - 1
    xa = 0.5 * (om.mesh.V[A[:, 0], 0] + om.mesh.V[A[:, 1], 0])
    ya = 0.5 * (om.mesh.V[A[:, 0], 1] + om.mesh.V[A[:, 1], 1])
    vel = {"U_SSA": 100 * np.sin(x / 3e5) + 20, "V_SSA": 50 * np.cos(y / 2e5),
           "Ux_SSA_Ac": 100 * np.sin(xa / 3e5) + 20, "Uy_SSA_Ac": 50 * np.cos(ya / 2e5)}
    return wl, om, ice, vel


def test_picard_rn_is_the_relative_l2_change_of_n():
    """RN = sqrt(sum (N - N_prev)^2 / sum N^2) (src/ice_dynamics_module.f90:512-516):
    with N_prev = N/2 it is exactly 1/2 (without the root 1/4, with the ratio
    inverted 1); with a random N_prev it equals the numpy value."""
    wl, om, ice, vel = _rn_state()
    p = wl.params
    p.SSA_max_outer_loops, p.SSA_max_inner_loops, p.SSA_RN_tol = 1, 1, 0.0
    ice.set(**vel)
    ice["N_AaAc"][:] = 0.0
    ice.solve_SSA(p)
    N1 = ice["N_AaAc"].copy()
    assert np.all(N1 > 0)
    ice.set(**vel)
    ice["N_AaAc"][:] = 0.5 * N1
    st = ice.solve_SSA(p)
    assert np.array_equal(ice["N_AaAc"], N1)        # same velocities, same viscosity
    assert abs(st.last_RN - 0.5) < 1e-14, st.last_RN
    rng = np.random.default_rng(4)
    Nr = N1 * np.exp(0.3 * rng.standard_normal(N1.size))
    ice.set(**vel)
    ice["N_AaAc"][:] = Nr
    st = ice.solve_SSA(p)
    ref = np.sqrt(np.sum((N1 - Nr) ** 2) / np.sum(N1 ** 2))
    assert abs(st.last_RN - ref) < 1e-13 * ref
    # converged on RN < SSA_RN_tol: no sweep is taken
    ice.set(**vel)
    ice["N_AaAc"][:] = 0.5 * N1
    p.SSA_RN_tol = 0.5000001
    st = ice.solve_SSA(p)
    assert st.picard_converged == 1 and st.total_sweeps == 0


def test_first_rn_is_one_and_failure_paths():
    wl, om, ice = _small_ssa()
    p = wl.params
    p.SSA_max_outer_loops = 1
    st = ice.solve_SSA(p)
    assert st.last_RN == 1.0                         # N starts at zero
    # max inner loops reached: warning status, not an error
    wl, om, ice = _small_ssa()
    p = wl.params
    p.SSA_max_outer_loops, p.SSA_max_inner_loops, p.SSA_max_residual_UV = 2, 1, 1e-12
    st = ice.solve_SSA(p)
    assert st.status == 1 and st.sweeps[0] == 1
    # divergent over-relaxation: the velocities are reset, and a second reset aborts
    wl, om, ice = _small_ssa()
    p = wl.params
    p.SSA_SOR_omega, p.SSA_max_outer_loops = 2.6, 5
    st = ice.solve_SSA(p)
    assert st.resets == 2 and st.status == 3
    # no grounded ice: velocities set to zero
    wl, om, ice = _small_ssa()
    ice["mask_sheet"][:] = 0
    ice["U_SSA"][:] = 5.0
    st = ice.solve_SSA(wl.params)
    assert st.picard_iterations == 0 and np.all(ice["U_SSA"] == 0.0)


def _mc_state(n=1500, seed=9):
    m = meshgen.generate_mesh(geometry.dome_resolution(), n, seed=seed)
    om = po.OracleMesh(m).build(colour=False)
    ice = po.OracleIce(om)
    return m, om, ice


def test_mass_conservation_and_positivity():
    m, om, ice = _mc_state()
    rng = np.random.default_rng(1)
    H = 500 + 500 * rng.random(m.nV)
    ice["Hi"][:] = H
    ice["Up_SIA_Ac"][:] = rng.standard_normal(om.nAc) * 50.0
    ice["Up_SSA_Ac"][:] = rng.standard_normal(om.nAc) * 50.0
    SMB = 0.3 * np.ones(m.nV)
    p = DynParams(zero_Hi_at_edges=False)
    ice.calculate_ice_thickness_change(SMB, np.zeros(m.nV), 2.0, np.zeros(m.nV, np.int32), p)
    before = np.sum(om["A"] * H) + np.sum(SMB * om["A"] * 2.0)
    assert np.isclose(np.sum(om["A"] * ice["Hi"]), before, rtol=1e-12)
    # strong outflow and melt: the limiter keeps H >= 0 and volume is still conserved
    m, om, ice = _mc_state(seed=10)
    H = 50 * rng.random(m.nV)
    ice["Hi"][:] = H
    ice["Up_SIA_Ac"][:] = rng.standard_normal(om.nAc) * 5e4
    ice["Up_SSA_Ac"][:] = 0.0
    melt = -10 * rng.random(m.nV)
    ice.calculate_ice_thickness_change(melt, np.zeros(m.nV), 1.0, np.zeros(m.nV, np.int32), p)
    Hn = ice["Hi"]
    assert Hn.min() > -1e-9 * H.max()
    removed = np.maximum(melt * om["A"], -om["A"] * H)       # melt limited by available ice
    assert np.isclose(np.sum(om["A"] * Hn), np.sum(om["A"] * H) + np.sum(removed), rtol=1e-9)


def test_uniform_translation_leaves_uniform_ice_unchanged():
    m, om, ice = _mc_state(seed=11)
    ice["Hi"][:] = 1000.0
    A = om["Aci"] - 1
    d = m.V[A[:, 1]] - m.V[A[:, 0]]
    L = np.hypot(d[:, 0], d[:, 1])
    U = np.array([120.0, -45.0])
    ice["Up_SIA_Ac"][:] = (U[0] * d[:, 0] + U[1] * d[:, 1]) / L
    ice["Up_SSA_Ac"][:] = 0.0
    p = DynParams(zero_Hi_at_edges=False)
    ice.calculate_ice_thickness_change(np.zeros(m.nV), np.zeros(m.nV), 1.0, np.zeros(m.nV, np.int32), p)
    inner = m.edge_index == 0
    scale = 1000.0 * np.hypot(*U) * np.sqrt(om["A"])     # flux magnitude per unit area scale
    assert np.all(np.abs(ice["dHi_dt"][inner]) * om["A"][inner] <= 1e-8 * scale[inner] * np.sqrt(om["A"][inner]))


def test_zero_timestep_and_edges():
    m, om, ice = _mc_state(seed=12)
    ice["Hi"][:] = 100.0
    ice["Up_SIA_Ac"][:] = 10.0
    ice.calculate_ice_thickness_change(np.ones(m.nV), np.zeros(m.nV), 0.0, np.zeros(m.nV, np.int32), DynParams())
    assert np.all(ice["dHi_dt"] == 0.0)
    assert np.all(ice["Hi"][m.edge_index > 0] == 0.0) and np.all(ice["Hi"][m.edge_index == 0] == 100.0)
    noice = (np.arange(m.nV) % 7 == 0).astype(np.int32)
    ice["Hi"][:] = 100.0
    ice.calculate_ice_thickness_change(np.zeros(m.nV), np.zeros(m.nV), 1.0, noice, DynParams())
    assert np.all(ice["Hi"][noice == 1] == 0.0)


def _gl_setup(H_scale=1.0, A=1e-16):
    """Marine sheet on a deep bed (Hb < -1000 m near the GL, so the friction angle
    sits at its 5 degree floor): general update and yield stress on the oracle."""
    from ufemism_amd.dynamics import initial_fields
    wl = geometry.marine(3000, seed=21, picard=2)
    r = np.hypot(wl.mesh.V[:, 0], wl.mesh.V[:, 1])
    wl.Hb = -1500.0 - 1.5 * r / 1e3
    om = po.OracleMesh(wl.mesh).build()
    ice = po.OracleIce(om)
    for k, v in initial_fields(wl, om["VAaAc"], wl.params.nZ).items():
        if k in ice.arr:
            ice[k][...] = v
    ice["Hi"][...] = H_scale * ice["Hi"]
    ice["Hb"][...] = H_scale * ice["Hb"]                 # same flotation crossing at every scale
    ice["A_flow_mean"][...] = A
    ice.update_general_ice_model_data()
    lib = po.load()
    lib.om_basal_yield_stress(ct.byref(om.s), ct.byref(ice.s))
    lib.om_calculate_GL_flux(ct.byref(om.s), ct.byref(ice.s))
    return wl, om, ice


def test_gl_flux_scaling_and_direction():
    """calculate_GL_flux (src/ice_dynamics_module.f90:848-949): the Tsai (2015)
    flux scales as A H^(n+2) (doubling thickness and depth keeps the GL and the
    5 degree friction angle, and multiplies the flux by 2^5), points down the
    thickness-above-flotation gradient, and Qp is its projection on vi -> vj."""
    wl, om, ice = _gl_setup()
    gl = ice["mask_gl_Ac"] == 1
    assert gl.sum() > 20
    assert np.allclose(ice["phi_fric_AaAc"][: wl.mesh.nV][ice["mask_sheet"] == 1], 5.0)
    Q = ice["Qabs_GL_Ac"].copy()
    assert np.all(Q[gl] > 0) and np.all(Q[~gl] == 0)
    _, _, ice2 = _gl_setup(H_scale=2.0)
    assert np.allclose(ice2["Qabs_GL_Ac"][gl], 32.0 * Q[gl], rtol=1e-12)
    _, _, ice3 = _gl_setup(A=3e-16)
    assert np.allclose(ice3["Qabs_GL_Ac"][gl], 3.0 * Q[gl], rtol=1e-12)
    # direction: antiparallel to grad(TAF) on the Ac mesh
    k = 1028.0 / 910.0
    gx = ice["dHi_dx_Ac"] - (ice["dSL_dx_Ac"] - ice["dHb_dx_Ac"]) * k
    gy = ice["dHi_dy_Ac"] - (ice["dSL_dy_Ac"] - ice["dHb_dy_Ac"]) * k
    ux, uy = ice["Ux_SSA_Ac"][gl], ice["Uy_SSA_Ac"][gl]
    gn = np.hypot(gx[gl], gy[gl])
    un = np.hypot(ux, uy)
    assert np.allclose((ux * gx[gl] + uy * gy[gl]) / (un * gn), -1.0, atol=1e-12)
    # Qp: flux component along the Ac connection vi -> vj
    A = om["Aci"] - 1
    d = wl.mesh.V[A[:, 1]] - wl.mesh.V[A[:, 0]]
    d /= np.hypot(d[:, 0], d[:, 1])[:, None]
    cosang = (d[gl, 0] * ux + d[gl, 1] * uy) / un
    assert np.allclose(ice["Qp_GL_Ac"][gl], Q[gl] * cosang, rtol=1e-10, atol=1e-12 * Q.max())
    # the thickness at the crossing times the speed is the flux magnitude
    Hi, Hb, SL = ice["Hi"], ice["Hb"], ice["SL"]
    ti = Hi[A[gl, 0]] - (SL[A[gl, 0]] - Hb[A[gl, 0]]) * k
    tj = Hi[A[gl, 1]] - (SL[A[gl, 1]] - Hb[A[gl, 1]]) * k
    lam = ti / (ti - tj)
    assert np.all((lam > 0) & (lam < 1))
    HGL = Hi[A[gl, 0]] * (1 - lam) + Hi[A[gl, 1]] * lam
    assert np.allclose(un * HGL, Q[gl], rtol=1e-12)


def test_gl_flux_rows_are_prescribed_in_solve_SSA():
    """With the GL flux on, solve_SSA keeps the GL Ac velocities it prescribed
    (src/ice_dynamics_module.f90:624-631 skips them in every sweep).  A small
    rate factor keeps the prescribed speeds moderate: a reset (:679-684) would
    zero them like every other velocity."""
    wl, om, ice = _gl_setup(A=1e-21)
    ice["A_flow_mean_Ac"][...] = 1e-21
    gl = ice["mask_gl_Ac"] == 1
    ux0, uy0 = ice["Ux_SSA_Ac"][gl].copy(), ice["Uy_SSA_Ac"][gl].copy()
    assert 10.0 < np.abs(ux0).max() < 1e5
    p = DynParams(m_enh_sia=1.0, m_enh_ssa=1.0, SSA_max_outer_loops=3, use_tau_c_override=True,
                  use_analytical_GL_flux=True)
    st = ice.solve_SSA(p)
    assert st.picard_iterations == 3 and st.resets == 0
    assert np.array_equal(ice["Ux_SSA_Ac"][gl], ux0) and np.array_equal(ice["Uy_SSA_Ac"][gl], uy0)
    assert np.abs(ice["Ux_SSA_Ac"][~gl]).max() > 0


def test_gl_rows_are_fixed_rows_of_the_direct_solution():
    """With the GL flux on, SOR that skips the GL Ac rows (src/ice_dynamics_module.f90:
    624-631) converges to the direct solution of the system in which those rows
    are identity rows holding the prescribed values: the system that the Krylov
    solves (residual and correction zero on those rows) also solve.  Without
    the fixed rows the same solution is far from the direct one."""
    wl, om, ice = _gl_setup(A=1e-21)
    ice["A_flow_mean_Ac"][...] = 1e-21
    p = DynParams(m_enh_sia=1.0, m_enh_ssa=1.0, SSA_max_outer_loops=1, use_tau_c_override=True,
                  use_analytical_GL_flux=True, SSA_max_residual_UV=1e-10, SSA_max_inner_loops=400000)
    st = ice.solve_SSA(p)
    assert st.status == 0 and st.sweeps[0] < 400000 and st.resets == 0
    nV = om.mesh.nV
    fixed = np.zeros(om.nVAaAc, bool)
    fixed[nV:] = ice["mask_gl_Ac"] == 1
    edge = np.concatenate([om.mesh.edge_index > 0, om["edge_index_Ac"] > 0])
    assert (fixed & ~edge).sum() > 20
    err = _direct_error(om, ice, fixed)
    assert err < 1e-8, err
    assert _direct_error(om, ice) > 1e-4

