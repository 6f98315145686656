"""One UFEMISM dynamics step on the GPU, in the reference's call order.

The reference's time loop calls, per step (src/UFEMISM_main_model.f90:90-132):
calculate_ice_thickness_change, update_general_ice_model_data, solve_SIA and
solve_SSA.  ``Step.run`` performs update_general_ice_model_data -> solve_SIA
-> solve_SSA -> calculate_ice_thickness_change on a freshly initialised
state, which is the single-step workload of the benchmark configurations.
"""
from __future__ import annotations

import dataclasses

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


def frozen_ssa_system(ctx: abi.Context, wl, picard=3, warm_tol=1e-6):
    """Set up a workload and freeze one linearised SSA system at a flowing
    state: general update, SIA, then `picard` Picard steps of solve_SSA whose
    linear systems are solved to the relative residual `warm_tol` (FP64
    Krylov, multigrid-preconditioned), and finally ufe_ssa_prepare, which
    rebuilds viscosity, sliding and coefficients from those velocities (the
    next Picard step's system, src/ice_dynamics_module.f90:501-531).  The
    reference's own SOR rule (max |dU| < 2.5 m/yr) stops after one sweep on
    fine meshes from rest, so it would leave the velocities near zero.  The
    parameters are left at the workload's FP64 reference setting."""
    import copy
    step = Step(ctx, wl)
    ctx.update_general_ice_model_data()
    ctx.solve_SIA()
    p = copy.deepcopy(wl.params)
    p.SSA_max_outer_loops = picard
    p.SSA_RN_tol = 0.0
    ctx.set_params(p, abi.UFE_FP64_KRYLOV, linear_rel_tol=warm_tol, gmres_restart=20,
                   precond_sweeps=2, precond_type=abi.UFE_PRECOND_MULTIGRID)
    st = ctx.solve_SSA()
    ctx.set_params(wl.params, abi.UFE_FP64)
    ctx.ssa_prepare()
    return step, st


@dataclasses.dataclass
class TimeConfig:
    """Time-step configuration, src/configuration_module.f90:34-44 (defaults)."""
    start_time_of_run: float = 0.0
    dt_max: float = 10.0
    dt_thermo: float = 10.0
    dt_climate: float = 10.0
    dt_SMB: float = 10.0
    dt_BMB: float = 10.0
    dt_output: float = 5000.0
    dt_bedrock_ELRA: float = 100.0


_ACTIONS = ("SIA", "SSA", "thermo", "climate", "SMB", "BMB", "ELRA", "output")


class RegionClock:
    """The time bookkeeping of a model region: the start values of
    initialise_model_region (src/UFEMISM_main_model.f90:351-396) and
    determine_timesteps_and_actions (:708-843), given the critical time steps
    that ufe_critical_timesteps computes on the device (:734-780)."""

    def __init__(self, cfg: TimeConfig = None):
        self.cfg = cfg = cfg or TimeConfig()
        t = cfg.start_time_of_run
        self.time = t
        self.t0 = {a: t for a in _ACTIONS}
        self.t1 = {a: t for a in _ACTIONS}
        self.t1["thermo"] = t + cfg.dt_thermo
        self.dt = 0.0
        self.dt_prev = 1000.0
        self.dt_SIA = self.dt_SSA = 0.0
        self.do = {a: True for a in _ACTIONS}
        self.do["thermo"] = False
        self.critical = None

    def determine_timesteps_and_actions(self, dt_crit, t_end):
        """dt_crit: (dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min), already
        multiplied by the 0.9 correction factor.  Sets dt, the do-flags and
        advances time, as the reference does (:782-839)."""
        cfg = self.cfg
        dt_D, dt_S, dt_3 = (float(x) for x in dt_crit)
        self.critical = min(min(min(dt_D, dt_S), dt_3), cfg.dt_max)
        if abs(1.0 - self.critical / self.dt_prev) > 0.1:
            self.dt_prev = self.critical
        self.dt_SIA = min(cfg.dt_max, min(dt_D, dt_3))
        self.dt_SSA = min(cfg.dt_max, dt_S)
        step = {"SIA": self.dt_SIA, "SSA": self.dt_SSA, "thermo": cfg.dt_thermo, "climate": cfg.dt_climate,
                "SMB": cfg.dt_SMB, "BMB": cfg.dt_BMB, "ELRA": cfg.dt_bedrock_ELRA, "output": cfg.dt_output}
        for a in _ACTIONS:
            self.t1[a] = self.t0[a] + step[a]
        t_next = min(self.t1[a] for a in _ACTIONS)
        self.dt = t_next - self.time
        self.do = {a: t_next == self.t1[a] for a in _ACTIONS}
        if t_next >= t_end:
            # crop the step to the coupling interval; everything but output runs
            self.dt = t_end - self.time
            for a in ("SIA", "SSA", "thermo", "climate", "SMB", "BMB"):
                self.do[a] = True
        self.time = self.time + self.dt


def run_dynamics(ctx: abi.Context, t_end, clock: RegionClock, dt_first=0.0, on_step=None):
    """The ice-dynamics part of run_model's time loop (src/UFEMISM_main_model.f90:74-215)
    on the device-resident state: calculate_ice_thickness_change with the
    previous step, update_general_ice_model_data, solve_SIA and solve_SSA when
    their time has come, solve_SIA_3D with the thermodynamics interval (the
    velocity part of update_ice_temperature), then the critical time steps and
    the next actions.  Climate, SMB, BMB, bedrock and output are not modelled
    here: SMB_year, BMB and the bed stay as set, and their clocks only advance.
    Returns the list of (time, dt, do-flags, SSA stats or None) per step."""
    log = []
    clock.dt = dt_first
    while clock.time < t_end:
        ctx.calculate_ice_thickness_change(clock.dt)
        ctx.update_general_ice_model_data()
        st = None
        if clock.do["SIA"]:
            ctx.solve_SIA()
            clock.t0["SIA"] = clock.time
        if clock.do["SSA"]:
            st = ctx.solve_SSA()
            clock.t0["SSA"] = clock.time
        for a in ("climate", "SMB", "BMB"):
            if clock.do[a]:
                clock.t0[a] = clock.time
        if clock.do["thermo"]:
            ctx.solve_SIA_3D()
            clock.t0["thermo"] = clock.time
        if clock.do["output"]:
            clock.t0["output"] = clock.time
        if clock.do["ELRA"]:
            clock.t0["ELRA"] = clock.time
        clock.determine_timesteps_and_actions(ctx.critical_timesteps(), t_end)
        log.append((clock.time, clock.dt, dict(clock.do), st))
        if on_step:
            on_step(clock)
    return log
