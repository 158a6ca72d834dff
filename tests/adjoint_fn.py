"""Test-only torch.autograd.Function over the C ABI: z -> u = K(E0 exp(z))^-1 f.

Forward: device assembly (svfsi_assemble_dev with per-element modulus) and
the FP64 device solve.  Backward: transposed device solve
(fsils_gpu_solve_t_dev) and the element dK/dE . u kernel
(svfsi_lelas_dkde_dot_dev), chained with dE/dz = E."""
import torch

from svfsi_amd.adjoint import LelasAdjoint


class LelasSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, z, adj: LelasAdjoint, E0: float):
        E = E0 * torch.exp(z)
        u, _ = adj.forward(E.contiguous())
        ctx.save_for_backward(E, u)
        ctx.adj = adj
        return u

    @staticmethod
    def backward(ctx, gu):
        E, u = ctx.saved_tensors
        gE, _, _ = ctx.adj.grad_E(u, gu)
        return gE * E, None, None


class FluidViscositySolve(torch.autograd.Function):
    """z -> du = K(mu)^-1 R(mu), mu = mu0 exp(z) (one linearised VMS step)."""

    @staticmethod
    def forward(ctx, z, adj, mu0: float):
        mu = (mu0 * torch.exp(z)).contiguous()
        du, _ = adj.forward(mu)
        ctx.save_for_backward(mu, du)
        ctx.adj = adj
        return du

    @staticmethod
    def backward(ctx, gdu):
        mu, du = ctx.saved_tensors
        g, _, _ = ctx.adj.grad_mu(mu, du, gdu)
        return g * mu, None, None
