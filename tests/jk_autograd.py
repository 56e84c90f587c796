"""Check-harness wrapper (tests only): the two-electron energy E = 1/2 tr(DJ) - 1/4 tr(DK) as a
differentiable function of the atomic coordinates at fixed density, for torch.autograd.

forward runs lbx_jk (through lible_amd.lbx) on a context built at the given coordinates;
backward runs the library's adjoint gradient kernel (lbx_jk_grad) on the same context and
scales it by the incoming gradient.  The reference has no such operator: its derivative surface
is the per-quartet ERI4D1Kernel (eri_kernels.hpp:113-131), which the gradient kernel contracts
with the density on the fly.
"""
from __future__ import annotations

import torch

from lible_amd import basis as B
from lible_amd import lbx


class JKEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coords: torch.Tensor, D: torch.Tensor, shells, schwarz: float):
        mol = B.Molecule(coords.detach().cpu().numpy().astype("float64"), shells)
        c = lbx.Context(mol)
        Dd = D.to(device=c.device, dtype=torch.float64).contiguous()
        J, K, _ = c.jk(Dd, schwarz=schwarz)
        E = 0.5 * (Dd * J).sum() - 0.25 * (Dd * K).sum()
        ctx.lbx_ctx = c
        ctx.D = Dd
        ctx.schwarz = schwarz
        return E.to(coords.device)

    @staticmethod
    def backward(ctx, grad_E):
        g, _ = ctx.lbx_ctx.jk_grad(ctx.D, schwarz=ctx.schwarz)
        return grad_E * g.to(grad_E.device), None, None, None


def jk_energy(coords: torch.Tensor, D: torch.Tensor, shells, schwarz: float = 1e-12) -> torch.Tensor:
    """E(coords) at fixed D; coords (natoms, 3) float64 in bohr, requires_grad for a gradient."""
    return JKEnergy.apply(coords, D, shells, schwarz)
