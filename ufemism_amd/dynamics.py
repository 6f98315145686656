"""One UFEMISM dynamics step on the GPU, in the reference's call order.

The reference's time loop calls, per step (src/UFEMISM_main_model.f90:90-132):
calculate_ice_thickness_change, update_general_ice_model_data, solve_SIA and
solve_SSA.  ``Step.run`` performs update_general_ice_model_data -> solve_SIA
-> solve_SSA -> calculate_ice_thickness_change on a freshly initialised
state, which is the single-step workload of the benchmark configurations.
"""
from __future__ import annotations

import numpy as np

from . import abi

OUTPUTS = ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "Up_SSA_Ac", "Uo_SSA_Ac",
           "U_SIA", "V_SIA", "D_SIA", "Up_SIA_Ac", "Uo_SIA_Ac", "D_SIA_Ac",
           "Hi", "dHi_dt", "Hs", "dHs_dx", "dHs_dy", "dHs_dx_shelf", "dHs_dy_shelf",
           "mask", "mask_Ac", "eta_AaAc", "N_AaAc", "tau_c_AaAc", "S_AaAc")


def initial_fields(wl, VAaAc, nZ):
    """Initial ice state of a workload, keyed by reference field name."""
    nV = wl.mesh.nV
    extra = {"Ti": wl.temperature()} if wl.params.physical_properties == 1 else {}
    return extra | {
        "Hi": wl.Hi, "Hb": wl.Hb, "SL": wl.SL,
        "A_flow": np.full((nV, nZ), wl.A_flow), "A_flow_mean": np.full(nV, wl.A_flow),
        "A_flow_Ac": wl.A_flow, "A_flow_mean_Ac": wl.A_flow,
        "tau_c_override_AaAc": wl.tau_c_override(VAaAc),
        "SMB_year": wl.SMB, "BMB": wl.BMB,
    }


class Step:
    """Mesh set-up plus one dynamics step of a workload on one context."""

    def __init__(self, ctx: abi.Context, wl, sweep_batch=0):
        self.ctx, self.wl = ctx, wl
        ctx.set_params(wl.params, abi.UFE_FP64, sweep_batch)
        self.ma = ctx.mesh_build(wl.mesh)
        ctx.mesh_upload(self.ma)
        self.init = initial_fields(wl, self.ma["VAaAc"], wl.params.nZ)
        self.reset()

    def reset(self):
        """Back to the initial state: input fields, and the SSA at rest (the
        velocities and the Picard viscosity that solve_SSA warm-starts from)."""
        for k, v in self.init.items():
            self.ctx.set(k, v)
        for k in ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac", "N_AaAc"):
            self.ctx.fill(k, 0.0)

    def run(self):
        c = self.ctx
        c.update_general_ice_model_data()
        c.solve_SIA()
        st = c.solve_SSA()
        c.calculate_ice_thickness_change(self.wl.dt)
        out = {k: c.get(k) for k in OUTPUTS}
        out["sweeps"] = list(st.sweeps)[: st.picard_iterations]
        out["picard"] = st.picard_iterations
        out["stats"] = st
        return out
