"""torch.autograd harness of the discrete adjoint (semo_mesh_amd.autograd).

The Function's backward is semo_adjoint_alpha_p.  Checked here:
  * J and alpha.grad from the autograd graph equal the adjoint entry point's
    own (J, dJ/dalpha) bit for bit, on a CUDA tensor;
  * one solver serves repeated calls: after a reset the second call's J and
    gradient equal the first's;
  * torch.autograd.gradcheck: the backward against torch's own central
    differences of the forward, every component, on a 4^3 jittered cavity
    (step 1e-4, solves at 1e-14; relative 1e-5 of the largest component).
Case: rho = 1, mu = 0.01, dt = 0.05, alpha_p ~ U[0, 10) seeded, 20 outer
iterations of one time step, J = int p dA (zmax) - int p dA (zmin).
"""
import numpy as np
import pytest
import torch

from semo_mesh_amd.cases import cavity_case
from semo_mesh_amd.mesh import cavity_mesh

pytestmark = pytest.mark.gpu

N_ITER = 20
W = {"zmax": 1.0, "zmin": -1.0}


def _factory(mesh):
    from semo_mesh_amd.solver import GpuSolver
    case = cavity_case(dt=0.05)
    case.rho = (1.0, 1.0)
    case.mu = (0.01, 0.01)

    def make():
        g = GpuSolver(0)
        g.upload_mesh(mesh, reorder=True)
        g.set_case(case)
        g.set_solver("fp64", tol_p=1e-14, tol_u=1e-14, maxit_p=300, maxit_u=1000, coarse_size=1000)
        return g
    return make


def _reset(g):
    # the cavity at rest: the state a fresh solver starts from
    n = g.mesh.n_cells
    for f in ("U", "V", "W", "P"):
        g.set_field(f, np.zeros(n))


def test_autograd_matches_adjoint_entry_point():
    from semo_mesh_amd.autograd import pressure_objective
    mesh = cavity_mesh(8, jitter=0.1, seed=11)
    make = _factory(mesh)
    a0 = 10.0 * np.random.default_rng(5).random(mesh.n_cells)
    alpha = torch.tensor(a0, dtype=torch.float64, device="cuda", requires_grad=True)
    solver = make()
    try:
        J = pressure_objective(alpha, solver, N_ITER, W, _reset)
        J.backward()
        g = make()
        g.set_field("alphaP", a0)
        J_ref, grad_ref = g.adjointAlphaP(N_ITER, W)
        g.close()
        assert J.device.type == "cuda" and alpha.grad.device.type == "cuda"
        assert J.item() == J_ref
        assert np.array_equal(alpha.grad.cpu().numpy(), grad_ref)
        # the same solver again, from the reset state
        alpha.grad = None
        J2 = pressure_objective(alpha, solver, N_ITER, W, _reset)
        J2.backward()
        assert abs(J2.item() - J_ref) <= 1e-12 * abs(J_ref)
        assert np.abs(alpha.grad.cpu().numpy() - grad_ref).max() <= 1e-10 * np.abs(grad_ref).max()
        # without a gradient: the plain iterations, the same J
        with torch.no_grad():
            J3 = pressure_objective(alpha.detach(), solver, N_ITER, W, _reset)
        assert abs(J3.item() - J_ref) <= 1e-12 * abs(J_ref)
    finally:
        solver.close()


def test_autograd_gradcheck():
    from semo_mesh_amd.autograd import pressure_objective
    mesh = cavity_mesh(4, jitter=0.1, seed=29)
    make = _factory(mesh)
    a0 = 10.0 * np.random.default_rng(31).random(mesh.n_cells)
    alpha = torch.tensor(a0, dtype=torch.float64, device="cuda", requires_grad=True)
    g = make()
    g.set_field("alphaP", a0)
    _, grad = g.adjointAlphaP(N_ITER, W)
    g.close()
    atol = 1e-5 * float(np.abs(grad).max())
    solver = make()
    try:
        assert torch.autograd.gradcheck(lambda a: pressure_objective(a, solver, N_ITER, W, _reset), (alpha,),
                                        eps=1e-4, atol=atol, rtol=1e-5)
    finally:
        solver.close()
