"""Adjoint gradients of a loss of the linear-system solution with respect to
per-element material parameters (no reference counterpart).

For K(theta) u = f with f independent of theta, and a loss L(u),

    K^T lam = dL/du,        dL/dtheta_e = -lam_e^T (dK_e/dtheta_e) u_e,

restricted to the free dofs: the solve's Dirichlet mask zeroes u and lam on
the constrained dofs, so only the free block of K enters.  The transposed
solve runs on the device (fsils_gpu_solve_t_dev, the forward solve's W K^T W
scaling); dK_e/dE_e . u is recomputed element by element
(svfsi_lelas_dkde_dot_dev) instead of storing element matrices.
"""
from __future__ import annotations

from . import fsils
from .problems import Problem
from .step import NewtonStep


class LelasAdjoint:
    """Forward solve and adjoint gradient of LELAS with per-element modulus
    E (device tensor, nEl).  Keeps the unscaled tangent of the last forward
    solve for the transposed solve."""

    def __init__(self, prob: Problem):
        self.prob = prob
        self.st = NewtonStep(prob)
        self.Val0 = None

    def forward(self, E, keep: bool = True):
        """u = K(E)^-1 f (FP64 device solve); returns (u, ls)."""
        st = self.st
        st.prop = E
        st.assemble()
        if keep:
            self.Val0 = st.Val.clone()
        ls = st.solve()
        if ls.RI.suc != 1:
            raise fsils.SvfsiError(f"forward solve did not converge ({ls.RI.itr} iterations)")
        return st.R.clone(), ls

    def residual(self, E, u):
        """r = f - K(E) u, with K(E) assembled and left unscaled in st.Val."""
        import torch

        st = self.st
        st.prop = E
        st.assemble()
        Ku = torch.empty_like(u)
        fsils.call("svfsi_spmv_dev", st.lhs.handle, self.prob.dof, fsils.dev_ptr(st.Val), fsils.dev_ptr(u),
                   fsils.dev_ptr(Ku), fsils.cur_stream())
        return st.R - Ku

    def solve_correction(self, E, u):
        """u(E) in correction form: u + K(E)^-1 (f - K(E) u).  The solve's
        relative tolerance then applies to the (small) correction."""
        st = self.st
        r = self.residual(E, u)
        st.R.copy_(r)
        ls = st.solve()
        if ls.RI.suc != 1:
            raise fsils.SvfsiError(f"correction solve did not converge ({ls.RI.itr} iterations)")
        return u + st.R

    def grad_E(self, u, dLdu):
        """dL/dE_e for the last forward solve (u, its tangent in Val0)."""
        import torch

        st = self.st
        if self.Val0 is None:
            raise fsils.SvfsiError("grad_E needs a forward solve first")
        lam = torch.empty_like(u)
        ls = st.make_ls()
        fsils.fsils_solve_t(st.lhs, ls, self.prob.dof, self.Val0, dLdu.contiguous(), lam)
        if ls.RI.suc != 1:
            raise fsils.SvfsiError(f"adjoint solve did not converge ({ls.RI.itr} iterations)")
        g = torch.empty(self.prob.mesh.nEl, dtype=torch.float64, device=u.device)
        st.dmesh.lelas_dkde_dot(self.prob.eq, lam, u, g)
        return g, lam, ls


class FluidViscosityAdjoint:
    """Linearised fluid step K(mu) du = R(mu) (FLUID3D VMS, per-element
    viscosity mu, device tensor nEl) and the adjoint gradient dL/dmu_e =
    lam_e^T (dR_e/dmu_e - dK_e/dmu_e du_e) with K^T lam = dL/du.  mixed
    selects the mixed-precision solver for the forward and transposed
    solves."""

    def __init__(self, prob: Problem, mixed: bool = False):
        self.prob = prob
        self.mixed = mixed
        self.st = NewtonStep(prob)
        self.Val0 = None

    def _solve(self, what: str):
        ls = self.st.solve(mixed=self.mixed)
        if ls.RI.suc != 1:
            raise fsils.SvfsiError(f"{what} did not converge ({ls.RI.itr} iterations)")
        return ls

    def forward(self, mu, keep: bool = True):
        st = self.st
        st.prop = mu
        st.assemble()
        if keep:
            self.Val0 = st.Val.clone()
        ls = self._solve("forward solve")
        return st.R.clone(), ls

    def solve_correction(self, mu, du):
        """du(mu) in correction form: du + K(mu)^-1 (R(mu) - K(mu) du)."""
        import torch

        st = self.st
        st.prop = mu
        st.assemble()
        Kx = torch.empty_like(du)
        fsils.call("svfsi_spmv_dev", st.lhs.handle, self.prob.dof, fsils.dev_ptr(st.Val), fsils.dev_ptr(du),
                   fsils.dev_ptr(Kx), fsils.cur_stream())
        st.R.sub_(Kx)
        self._solve("correction solve")
        return du + st.R

    def grad_mu(self, mu, du, dLdu):
        """dL/dmu_e for the last forward solve (its tangent kept in Val0)."""
        import torch

        st = self.st
        if self.Val0 is None:
            raise fsils.SvfsiError("grad_mu needs a forward solve first")
        lam = torch.empty_like(du)
        ls = st.make_ls()
        fsils.fsils_solve_t(st.lhs, ls, self.prob.dof, self.Val0, dLdu.contiguous(), lam, mixed=self.mixed)
        if ls.RI.suc != 1:
            raise fsils.SvfsiError(f"adjoint solve did not converge ({ls.RI.itr} iterations)")
        g = torch.empty(self.prob.mesh.nEl, dtype=torch.float64, device=du.device)
        st.dmesh.fluid_dmu_dot(self.prob.eq, st.Ag, st.Yg, mu, lam, du, g)
        return g, lam, ls
