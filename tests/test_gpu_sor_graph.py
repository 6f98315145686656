"""The reference FP64 SOR iteration of solve_SSA (src/ice_dynamics_module.f90:601-691)
with its sweep batches replayed from HIP graphs gives the same sweep counts
and the same bits as eager launches (UFE_SOR_GRAPH=0), and the graph path
still matches the oracle's step."""
import os

import numpy as np
import pytest

from ufemism_amd import abi, geometry
from ufemism_amd.dynamics import Step

pytestmark = pytest.mark.gpu


def _step(flag, wl):
    os.environ["UFE_SOR_GRAPH"] = flag
    ctx = abi.Context(0)
    try:
        out = Step(ctx, wl).run()
        return out
    finally:
        ctx.close()
        os.environ.pop("UFE_SOR_GRAPH", None)


@pytest.mark.parametrize("kind", ["dome8k", "marine20k"])
def test_graph_sweeps_equal_eager_sweeps(kind):
    print(f"[sor graph] {kind}", flush=True)
    wl = geometry.dome(8000, seed=1, picard=10) if kind == "dome8k" else geometry.marine(20000, seed=5, picard=10)
    a = _step("1", wl)
    b = _step("0", wl)
    print(f"  sweeps graph {a['sweeps']} eager {b['sweeps']}", flush=True)
    assert list(a["sweeps"]) == list(b["sweeps"])
    for k in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Hi"):
        assert np.array_equal(a[k], b[k]), k
    from oracle import pyoracle
    ref = pyoracle.run_step(wl)
    assert list(a["sweeps"]) == list(ref["sweeps"])
    for k in ("U_SSA", "V_SSA", "Hi"):
        e = np.linalg.norm(a[k] - ref[k]) / np.linalg.norm(ref[k])
        print(f"  {k}: rel L2 vs oracle {e:.2e}", flush=True)
        assert e <= 1e-9, (k, e)
