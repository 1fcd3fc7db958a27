"""torch.autograd binding of the discrete adjoint (no counterpart in the
reference, which has no adjoint).

``pressure_objective(alpha_p, solver, n_iter, weights, reset)`` returns

    J(alpha_p) = sum over patches of weight * sum_f A_f p_f

after ``n_iter`` outer iterations of one time step
(``incompressiblePressureBased(n_iter)``), started from the state that
``reset(solver)`` restores, with the Darcy porosity source -alpha_p U V in
the momentum equation.  Its backward is the hand-written discrete adjoint
(``semo_adjoint_alpha_p``): checkpointed forward iterates, transposed
face-flux Jacobians and transposed AMG-preconditioned solves on the GPU.

``solver`` is a configured ``GpuSolver`` (mesh, case and solver settings)
that lives across calls, so its mesh, geometry and AMG hierarchy are set up
once; ``reset`` puts every field of the initial state back (for a cavity
started from rest: U, V, W and P set to zero).  When ``alpha_p`` requires a
gradient, the forward runs ``semo_adjoint_alpha_p`` once -- the checkpointed
forward sweep that gives J, then the reverse sweep -- and keeps dJ/dalpha_p
on ``alpha_p``'s device for the backward, which only scales it: one library
call per gradient.  Without a gradient the forward runs the plain
iterations.  ``alpha_p`` crosses to the host once per call (the C ABI takes
host arrays); the gradient is returned on ``alpha_p``'s device.

Import torch before the first ``GpuSolver`` of the process: torch's CUDA
initialisation has failed ("No HIP GPUs are available") in a process that
had loaded libsemo_gpu.so and used the GPU through it first.
"""
from __future__ import annotations

import numpy as np
import torch


class PressureObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha_p: torch.Tensor, solver, n_iter: int, weights: dict, reset):
        a = alpha_p.detach().to("cpu", torch.float64).numpy()
        reset(solver)
        solver.set_field("alphaP", a)
        if ctx.needs_input_grad[0]:
            J, dJ = solver.adjointAlphaP(n_iter, weights)
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(dJ)).to(alpha_p.device))
            ctx.alpha_dtype = alpha_p.dtype
        else:
            solver.incompressiblePressureBased(n_iter)
            J = solver.objective(weights)
        return torch.tensor(J, dtype=torch.float64, device=alpha_p.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (dJ,) = ctx.saved_tensors
        grad = dJ * grad_out.to(torch.float64)
        return grad.to(ctx.alpha_dtype), None, None, None, None


def pressure_objective(alpha_p: torch.Tensor, solver, n_iter: int, weights: dict, reset) -> torch.Tensor:
    """J(alpha_p) as a differentiable scalar tensor (see PressureObjective)."""
    return PressureObjective.apply(alpha_p, solver, n_iter, weights, reset)
