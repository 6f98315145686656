"""Differentiable SSA velocities on the GPU.

``SSAVelocities`` is a ``torch.autograd.Function`` whose forward runs K Picard
steps of solve_SSA (src/ice_dynamics_module.f90:498-545) with tight linear
solves, and whose backward is the library's hand-written adjoint
(``ufe_ssa_adjoint_forward`` / ``ufe_ssa_adjoint_backward``).  Inputs and outputs
are float64 CUDA tensors over the AaAc points in the reference order (Aa
vertices 1..nV, then Ac points):

* ``log_tau_c``  log of the basal yield stress tau_c (Pa), the friction field;
* ``A_flow``     depth-mean rate factor A (Pa^-3 yr^-1);
* returns ``(U, V)``, the AaAc velocities after K Picard steps.

The gradient is exact for the K-step map with exact linear solves; the
linear solves stop at a relative residual ``lin_rel_tol`` (default 1e-12).
"""
from __future__ import annotations

import dataclasses

import torch

from . import abi

_VELOCITY_FIELDS = ("U_SSA", "V_SSA", "Ux_SSA_Ac", "Uy_SSA_Ac")


@dataclasses.dataclass
class AdjointOptions:
    picard_iterations: int = 10
    lin_rel_tol: float = 1e-12
    checkpoint_every: int = 1
    precision: int = abi.UFE_FP64
    gmres_restart: int = 40
    precond_sweeps: int = 5
    omega: float = 0.0
    max_sweeps: int = 1_000_000
    precond_type: int = abi.UFE_PRECOND_SOR   # UFE_PRECOND_MULTIGRID: both directions multigrid-preconditioned


class SSAProblem:
    """An uploaded mesh and ice state (after update_general_ice_model_data)
    whose SSA velocities are differentiated.  Every forward starts from the
    velocities present when the problem was created."""

    def __init__(self, ctx: abi.Context, opts: AdjointOptions | None = None):
        if not getattr(ctx, "_params", None) or not ctx._params.use_tau_c_override:
            raise ValueError("SSAProblem needs params with use_tau_c_override set")
        self.ctx = ctx
        self.opts = opts or AdjointOptions()
        self.nV = ctx.sizes["nV"]
        self.nAc = ctx.sizes["nAc"]
        self.nA = ctx.sizes["nVAaAc"]
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self._u0 = {}
        for name in _VELOCITY_FIELDS:
            n = self.nV if not name.endswith("_Ac") else self.nAc
            t = torch.empty(n, dtype=torch.float64, device=self.dev)
            ctx.get_device(name, t.data_ptr())
            self._u0[name] = t
        self.forward_stats = None
        self.backward_stats = None
        self.generation = 0              # forwards run; the device holds the last one's checkpoints

    def _set_inputs(self, log_tau_c: torch.Tensor, A_flow: torch.Tensor):
        for t in (log_tau_c, A_flow):
            if t.dtype != torch.float64 or t.device.type != "cuda" or t.numel() != self.nA:
                raise ValueError(f"expected a float64 CUDA tensor of {self.nA} AaAc values")
        tau = torch.exp(log_tau_c.detach()).contiguous()
        A = A_flow.detach().contiguous()
        torch.cuda.current_stream().synchronize()
        c = self.ctx
        c.set_device("tau_c_override_AaAc", tau.data_ptr())
        c.set_device("A_flow_mean", A[: self.nV].data_ptr())
        c.set_device("A_flow_mean_Ac", A[self.nV:].data_ptr())
        for name, t in self._u0.items():
            c.set_device(name, t.data_ptr())

    def forward(self, log_tau_c: torch.Tensor, A_flow: torch.Tensor):
        """Velocities (U, V) on the AaAc points; no graph."""
        self._set_inputs(log_tau_c, A_flow)
        o = self.opts
        self.generation += 1
        self.forward_stats = self.ctx.adjoint_forward(
            o.picard_iterations, o.lin_rel_tol, o.checkpoint_every, o.precision, o.gmres_restart,
            o.precond_sweeps, o.omega, o.max_sweeps, precond_type=o.precond_type)
        U = torch.empty(self.nA, dtype=torch.float64, device=self.dev)
        V = torch.empty_like(U)
        self.ctx.get_device("U_SSA_AaAc", U.data_ptr())
        self.ctx.get_device("V_SSA_AaAc", V.data_ptr())
        return U, V

    def backward(self, gU: torch.Tensor, gV: torch.Tensor, generation: int | None = None):
        """dJ/dlog(tau_c), dJ/dA for the last forward, given dJ/dU, dJ/dV."""
        if generation is not None and generation != self.generation:
            raise RuntimeError("the SSA checkpoints belong to a later forward; run the forward again")
        gU = gU.detach().to(torch.float64).contiguous()
        gV = gV.detach().to(torch.float64).contiguous()
        g_tau = torch.empty(self.nA, dtype=torch.float64, device=self.dev)
        g_A = torch.empty_like(g_tau)
        torch.cuda.current_stream().synchronize()
        self.backward_stats = self.ctx.adjoint_backward(gU.data_ptr(), gV.data_ptr(), g_tau.data_ptr(),
                                                        g_A.data_ptr())
        return g_tau, g_A


class SSAVelocities(torch.autograd.Function):
    """(log_tau_c, A_flow) -> (U, V) with the hand-written adjoint as backward."""

    @staticmethod
    def forward(fctx, log_tau_c, A_flow, problem: SSAProblem):
        fctx.problem = problem
        out = problem.forward(log_tau_c, A_flow)
        fctx.generation = problem.generation
        return out

    @staticmethod
    def backward(fctx, gU, gV):
        if gU is None:
            gU = torch.zeros(fctx.problem.nA, dtype=torch.float64, device=fctx.problem.dev)
        if gV is None:
            gV = torch.zeros_like(gU)
        g_tau, g_A = fctx.problem.backward(gU, gV, fctx.generation)
        return g_tau, g_A, None


def ssa_velocities(log_tau_c, A_flow, problem: SSAProblem):
    return SSAVelocities.apply(log_tau_c, A_flow, problem)
