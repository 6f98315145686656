"""The Fortran host (ice_dynamics_module over ISO_C_BINDING) drives one
dynamics step on the GPU and matches the CPU oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from ufemism_amd import build, geometry
from ufemism_amd.config import tau_c_from_beta

pytestmark = pytest.mark.gpu


def write_input(path, wl):
    m, p = wl.mesh, wl.params
    with open(path, "wb") as f:
        np.array([m.nV, m.nTri, m.nC_mem, p.nZ], np.int32).tofile(f)
        np.array([m.xmin, m.xmax, m.ymin, m.ymax], np.float64).tofile(f)
        for a in (m.V, m.nC, m.C, m.niTri, m.iTri, m.edge_index, m.Tri, m.A, m.Cw, m.Tri_edge_index):
            np.asfortranarray(a).ravel(order="F").tofile(f)
        np.asarray(p.zeta, np.float64).tofile(f)
        np.array([p.m_enh_sia, p.m_enh_ssa, p.SSA_RN_tol, p.SSA_max_residual_UV, p.SSA_SOR_omega, wl.dt],
                 np.float64).tofile(f)
        np.array([p.SSA_max_outer_loops, p.SSA_max_inner_loops], np.int32).tofile(f)
        for a in (wl.Hi, wl.Hb, wl.SL):
            np.asarray(a, np.float64).tofile(f)
        np.array([wl.A_flow, float(tau_c_from_beta(wl.beta))], np.float64).tofile(f)
        np.asarray(wl.SMB, np.float64).tofile(f)
        np.asarray(wl.BMB, np.float64).tofile(f)
        np.zeros(m.nV, np.int32).tofile(f)


def read_output(path, nV, nZ=15):
    with open(path, "rb") as f:
        head = np.fromfile(f, np.int32, 66)
        nAc, picard, sweeps = int(head[0]), int(head[1]), head[2:].tolist()
        vals = np.fromfile(f, np.float64)
    o, out = 0, {}
    for name, n in (("U_SSA", nV), ("V_SSA", nV), ("Up_SSA_Ac", nAc), ("Up_SIA_Ac", nAc),
                    ("Hi", nV), ("dHi_dt", nV), ("U_3D", nV * nZ), ("W_3D", nV * nZ), ("A", nV),
                    ("dt_crit", 3)):
        out[name] = vals[o:o + n]
        o += n
    for k in ("U_3D", "W_3D"):
        out[k] = out[k].reshape((nV, nZ), order="F")
    out["sweeps"] = sweeps[:picard]
    return out


def test_fortran_host_step_matches_oracle(tmp_path):
    exe = build.FORTRAN_EXE
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang)")
    from oracle import pyoracle
    print("[fortran] dome 3k step", flush=True)
    wl = geometry.dome(3000, seed=13, picard=6)
    write_input(tmp_path / "in.bin", wl)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = read_output(tmp_path / "out.bin", wl.mesh.nV)
    ref = pyoracle.run_step(wl)
    assert out["sweeps"] == ref["sweeps"]
    for k in ("U_SSA", "V_SSA", "Up_SSA_Ac", "Up_SIA_Ac", "Hi", "dHi_dt"):
        e = np.linalg.norm(out[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300)
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-9, k
    # the host computed the reference's Voronoi areas from Tri_edge_index
    assert np.array_equal(out["A"], ref["mesh"]["A"])
    # solve_SIA_3D and the critical time steps after the thickness change
    ice = ref["ice"]
    ice.solve_SIA_3D(wl.params)
    for k in ("U_3D", "W_3D"):
        e = np.linalg.norm(out[k] - ice[k]) / max(np.linalg.norm(ice[k]), 1e-300)
        print(f"  {k}: rel L2 {e:.2e}", flush=True)
        assert e <= 1e-9, k
    assert np.allclose(out["dt_crit"], ice.critical_timesteps(wl.params), rtol=1e-9)


def test_fortran_host_mixed_multigrid(tmp_path):
    """C%precision = UFE_MIXED and C%SSA_precond_type = multigrid through the
    Fortran configuration: the step equals the same step driven through the
    Python boundary with the same options."""
    exe = build.FORTRAN_EXE
    if not os.path.exists(exe):
        pytest.skip("Fortran host not built (flang)")
    from ufemism_amd import abi
    from ufemism_amd.dynamics import Step
    print("[fortran] dome 3k step, mixed + multigrid", flush=True)
    wl = geometry.dome(3000, seed=13, picard=6)
    write_input(tmp_path / "in.bin", wl)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "mixed_multigrid"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = read_output(tmp_path / "out.bin", wl.mesh.nV)
    ctx = abi.Context(0)
    try:
        step = Step(ctx, wl)
        ctx.set_params(wl.params, abi.UFE_MIXED, linear_rel_tol=1e-10, precond_type=abi.UFE_PRECOND_MULTIGRID)
        ref = step.run()
        print(f"  fortran sweeps {out['sweeps']} python sweeps {ref['sweeps']}", flush=True)
        for k in ("U_SSA", "V_SSA", "Hi"):
            e = np.linalg.norm(out[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300)
            print(f"  {k}: rel L2 {e:.2e}", flush=True)
            assert e <= 1e-7, k
        assert ctx.multigrid_info()["levels"] >= 2
    finally:
        ctx.close()
