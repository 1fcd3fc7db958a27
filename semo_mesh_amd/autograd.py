"""torch.autograd binding of the discrete adjoint (no counterpart in the
reference, which has no adjoint).

``PressureObjective.apply(alpha_p, make_solver, n_iter, weights)`` returns

    J(alpha_p) = sum over patches of weight * sum_f A_f p_f

after ``n_iter`` outer iterations of one time step
(``incompressiblePressureBased(n_iter)``), started from the state that
``make_solver()`` sets up, with the Darcy porosity source -alpha_p U V in the
momentum equation.  Its backward is the hand-written discrete adjoint
(``semo_adjoint_alpha_p``): checkpointed forward iterates, transposed
face-flux Jacobians and transposed AMG-preconditioned solves on the GPU.

``make_solver`` is a callable returning a configured ``GpuSolver`` (mesh,
case, solver settings and every field of the initial state set).  Forward
and backward each build their own solver from it, so both differentiate the
map from the same initial state; the backward's forward sweep is
bit-identical to the forward's.  ``alpha_p`` may live on any device; the
gradient is returned on that device in float64.
"""
from __future__ import annotations

import numpy as np
import torch


class PressureObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha_p: torch.Tensor, make_solver, n_iter: int, weights: dict):
        a = alpha_p.detach().to("cpu", torch.float64).numpy()
        g = make_solver()
        try:
            g.set_field("alphaP", a)
            g.incompressiblePressureBased(n_iter)
            J = g.objective(weights)
        finally:
            g.close()
        ctx.make_solver, ctx.n_iter, ctx.weights = make_solver, n_iter, dict(weights)
        ctx.save_for_backward(alpha_p.detach())
        return torch.tensor(J, dtype=torch.float64, device=alpha_p.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (alpha_p,) = ctx.saved_tensors
        a = alpha_p.to("cpu", torch.float64).numpy()
        g = ctx.make_solver()
        try:
            g.set_field("alphaP", a)
            _, dJ = g.adjointAlphaP(ctx.n_iter, ctx.weights)
        finally:
            g.close()
        grad = torch.from_numpy(np.ascontiguousarray(dJ)).to(alpha_p.device) * grad_out.to(torch.float64)
        return grad.to(alpha_p.dtype), None, None, None


def pressure_objective(alpha_p: torch.Tensor, make_solver, n_iter: int, weights: dict) -> torch.Tensor:
    """J(alpha_p) as a differentiable scalar tensor (see PressureObjective)."""
    return PressureObjective.apply(alpha_p, make_solver, n_iter, weights)
