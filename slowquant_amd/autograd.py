"""E(theta) = <psi0|U(theta)^+ H U(theta)|psi0> as a torch.autograd.Function.

forward runs the gate sweep psi = U(theta) psi0 and lam = H psi on the device
and keeps only those two vectors; backward is the hand-written adjoint
reverse sweep (sq_adjoint_reverse), which un-applies the gates on psi and lam
in place and returns dE/dtheta_i = 2 <lam_i|T_i psi_i> for every parameter
in one pass.  It replaces the theta block of the reference's
_calc_gradient_optimization (ups_wavefunction.py:1060-1142).
"""

from __future__ import annotations

import torch
from torch.autograd.function import once_differentiable

from . import engine as E


class UpsEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, program: E.Program, hamiltonian: E.Hamiltonian, psi0: torch.Tensor):
        if theta.dim() != 1 or theta.numel() != program.n_gates:
            raise ValueError(f"Expected {program.n_gates} theta values got {theta.numel()}")
        psi = program.state_prep(theta, psi0)
        lam, e = hamiltonian.sigma(psi)
        ctx.program = program
        ctx.theta = theta.detach().clone()
        ctx.vectors = (psi, lam)
        return torch.tensor(e, dtype=torch.float64, device=theta.device)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out: torch.Tensor):
        psi, lam = ctx.vectors
        ctx.vectors = None  # consumed by the reverse sweep
        g = ctx.program.adjoint_reverse(ctx.theta, psi, lam)
        grad = torch.as_tensor(g, dtype=torch.float64).to(ctx.theta.device)
        return grad_out * grad, None, None, None


def ups_energy(theta: torch.Tensor, program: E.Program, hamiltonian: E.Hamiltonian, psi0: torch.Tensor) -> torch.Tensor:
    """Differentiable energy (electronic + the Hamiltonian's core constant)."""
    return UpsEnergy.apply(theta, program, hamiltonian, psi0)
