"""Gradients of OperatorLike.integrate through a hand-written adjoint of the ETD step.

The reference differentiates integrate() with torch autograd through every elementwise op and FFT
of every step (_base.py:676-751 with integrator.step, _etdrk.py:72-82, _setdrk_step.py:14-25).
Here the forward pass runs the fused engine and keeps the spectral state every `every` steps
(checkpoints).  The backward pass walks the segments in reverse: it recomputes the states of one
segment from its checkpoint and applies tfsm_adjoint_step to each, from the last step back.
Memory: ceil(steps / every) + every states, instead of the reference's whole graph.

Supported: Convection in 1D, 2D and 3D (e.g. Burgers) and VorticityConvection in 2D (e.g.
NavierStokesVorticity, with explicit sources such as kolm_force), with ETDRK1/2, SETDRK1/2 and
SETDRK4 (the 'auto' scheme of nonlinear operators, _base.py:451-455); gradients with respect to the
initial field and to the coefficient of one linear term (a 0-d tensor, or one value per sample).
"""

import math
from typing import Optional

import torch

from . import _lib
from ._lib import check, lib
from .engine import SpectralPlan, _ptr, _stream

ADJOINT_SCHEMES = ("ETDRK1", "ETDRK2", "SETDRK1", "SETDRK2", "SETDRK4")


def _adjoint_workspace(plan: SpectralPlan) -> torch.Tensor:
    # the size depends on the plan's current scheme (SETDRK4 keeps more stages)
    nbytes = max(int(lib.tfsm_plan_adjoint_workspace_bytes(plan.handle)), 256)
    ws = getattr(plan, "_adj_ws", None)
    if ws is None or ws.numel() < nbytes:
        ws = torch.zeros(nbytes, dtype=torch.uint8, device=plan.device)
        plan._adj_ws = ws
    return ws


def adjoint_prepare(plan: SpectralPlan, grad_term: int):
    with torch.cuda.device(plan.device):
        check(lib.tfsm_adjoint_prepare(plan.handle, int(grad_term), _stream(plan.device)), "adjoint_prepare")


def adjoint_step(plan: SpectralPlan, u: torch.Tensor, ubar_next: torch.Tensor, ubar: torch.Tensor,
                 grad: Optional[torch.Tensor]):
    ws = plan.workspace()
    wsa = _adjoint_workspace(plan)
    with torch.cuda.device(plan.device):
        check(lib.tfsm_adjoint_step(plan.handle, _ptr(u), _ptr(ubar_next), _ptr(ubar), _ptr(grad), _ptr(ws), _ptr(wsa),
                                    _stream(plan.device)), "adjoint_step")


class IntegrateAdjoint(torch.autograd.Function):
    """u_T = Re ifftn(step^K(fftn(u_0))) with a hand-written backward.

    forward(u0, theta, plan, steps, every, grad_term): theta is the tensor coefficient of linear
    term grad_term (or a dummy tensor when grad_term < 0); it is not read in the forward pass (the
    plan already holds its value)."""

    @staticmethod
    def forward(ctx, u0, theta, plan: SpectralPlan, steps: int, every: int, grad_term: int):
        uh = plan.forward(u0.detach())
        ckpts = []
        s0 = 0
        while s0 < steps:
            n = min(every, steps - s0)
            ckpts.append((s0, n, uh.clone()))
            plan.step(uh, n)
            s0 += n
        ctx.plan, ctx.ckpts, ctx.grad_term = plan, ckpts, grad_term
        ctx.theta_shape = theta.shape
        ctx.theta_dtype = theta.dtype
        ctx.theta_device = theta.device
        return plan.inverse(uh)

    @staticmethod
    def backward(ctx, g):
        plan = ctx.plan
        npts = math.prod(plan.n)
        # adjoint of Re(ifftn(.)): (1 / prod N) * fftn(g)
        ub = plan.forward(g.detach().to(torch.float32).contiguous())
        ub.mul_(1.0 / npts)
        want_theta = ctx.grad_term >= 0 and ctx.needs_input_grad[1]
        adjoint_prepare(plan, ctx.grad_term if want_theta else -1)
        grad = torch.zeros(plan.batch, dtype=torch.float64, device=plan.device) if want_theta else None
        out = torch.empty_like(ub)
        for _, n, ck in reversed(ctx.ckpts):
            states = [ck.clone()]
            u = ck.clone()
            for _ in range(n - 1):
                plan.step(u, 1)
                states.append(u.clone())
            for i in range(n - 1, -1, -1):
                adjoint_step(plan, states[i], ub, out, grad)
                ub, out = out, ub
        g_u0 = plan.inverse(ub).mul_(float(npts)) if ctx.needs_input_grad[0] else None
        g_theta = None
        if want_theta:
            g = grad.sum() if math.prod(ctx.theta_shape) == 1 else grad
            g_theta = g.to(device=ctx.theta_device, dtype=ctx.theta_dtype).reshape(ctx.theta_shape)
        return g_u0, g_theta, None, None, None, None
