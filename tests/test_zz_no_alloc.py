"""The allocation-free FP64 SOR solve (ufe_ssa_sor_solve_noalloc) against
the plain FP64 SOR path of ufe_ssa_linear_solve, through the C-ABI: the same
sweeps, residuals and bit-identical velocities, on two calls in a row that
share the library's scratch array (a 20k marine system, then a smaller 5k
system on another context of the same GPU), and on a call stopped by its
sweep limit."""
import numpy as np
import pytest

from ufemism_amd import abi, geometry
from ufemism_amd.dynamics import Step

pytestmark = pytest.mark.gpu


def _system(ctx, wl, picard=4):
    Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    ctx.solve_SSA()
    ctx.ssa_prepare()


def _zero(ctx):
    ctx.fill("U_SSA_AaAc", 0.0)
    ctx.fill("V_SSA_AaAc", 0.0)


def _state(ctx):
    return ctx.get("U_SSA_AaAc"), ctx.get("V_SSA_AaAc")


def test_noalloc_sor_solve_equals_linear_solve():
    big, small = abi.Context(0), abi.Context(0)
    try:
        print("[noalloc] marine 20k and marine 5k systems", flush=True)
        _system(big, geometry.marine(20000, seed=5, picard=4))
        _system(small, geometry.marine(5000, seed=6, picard=4))
        assert small.sizes["nVAaAc"] < big.sizes["nVAaAc"]
        runs = {}
        # two allocation-free calls in a row, the second on the smaller system
        for name, c in (("big", big), ("small", small)):
            _zero(c)
            st = c.ssa_sor_solve_noalloc(1e-6, 100000, check_interval=16)
            runs[name] = (st, _state(c))
            print(f"  noalloc {name}: sweeps {st.sweeps} refinements {st.refinements} "
                  f"rel {st.rel_residual:.3e} status {st.status}", flush=True)
        for name, c in (("big", big), ("small", small)):
            _zero(c)
            ref = c.ssa_linear_solve(1e-6, 100000, precision=abi.UFE_FP64, check_interval=16)
            u, v = _state(c)
            st, (u1, v1) = runs[name]
            print(f"  linear_solve {name}: sweeps {ref.sweeps} rel {ref.rel_residual:.3e}", flush=True)
            assert st.status == ref.status == abi.UFE_OK
            assert st.sweeps == ref.sweeps and st.refinements == ref.refinements
            assert st.rel_residual == ref.rel_residual
            assert np.array_equal(u1, u) and np.array_equal(v1, v)
        # a call stopped by its sweep limit: same partial state and status
        _zero(small)
        st = small.ssa_sor_solve_noalloc(1e-12, 7, omega=1.1, check_interval=3)
        u1, v1 = _state(small)
        _zero(small)
        ref = small.ssa_linear_solve(1e-12, 7, precision=abi.UFE_FP64, omega=1.1, check_interval=3)
        u, v = _state(small)
        assert st.status == ref.status == abi.UFE_SOR_NOT_CONVERGED
        assert st.sweeps == ref.sweeps == 7 and st.rel_residual == ref.rel_residual
        assert np.array_equal(u1, u) and np.array_equal(v1, v)
    finally:
        big.close()
        small.close()


def test_noalloc_needs_a_mesh():
    c = abi.Context(0)
    try:
        st = abi.LinSolveStats()
        r = c.lib.ufe_ssa_sor_solve_noalloc(c.h, 1e-6, 10, 0.0, 16, __import__("ctypes").byref(st))
        assert r == abi.UFE_BAD_INPUT
    finally:
        c.close()
