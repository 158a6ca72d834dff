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


def _spmv(st: NewtonStep, Val, x, transpose: bool = False):
    """K x (or K^T x) with the block-CSR values Val on st's pattern."""
    import torch

    y = torch.empty_like(x)
    fsils.call("svfsi_spmv_t_dev" if transpose else "svfsi_spmv_dev", st.lhs.handle, st.prob.dof,
               fsils.dev_ptr(Val), fsils.dev_ptr(x), fsils.dev_ptr(y), fsils.cur_stream())
    return y


def _resid(st: NewtonStep, Val, x, b, transpose: bool = False):
    """b - K x (or b - K^T x) with double-double row sums, rounded once
    (svfsi_resid_dd_dev)."""
    import torch

    r = torch.empty_like(b)
    fsils.call("svfsi_resid_dd_dev", st.lhs.handle, st.prob.dof, fsils.dev_ptr(Val), fsils.dev_ptr(x),
               fsils.dev_ptr(b), fsils.dev_ptr(r), 1 if transpose else 0, fsils.cur_stream())
    return r


def _refined_correction(st: NewtonStep, Val, r, solve, passes: int, transpose: bool = False):
    """d with K d = r (K^T d = r), K the unscaled values Val, by `passes`
    solves in residual-correction form: d_0 = 0, d_{k+1} = d_k + K^-1 (r -
    K d_k).  Each residual is formed from the unscaled matrix with
    double-double row sums, so every pass gains the solver's relative
    tolerance times K's condition number, down to the working precision of
    d itself rather than to the FP64 residual floor eps |K| |d|.
    solve(rhs) returns K^-1 rhs to the solver's tolerance and leaves Val
    untouched."""
    import torch

    d = torch.zeros_like(r)
    rk = r
    for k in range(passes):
        d = d + solve(rk)
        if k + 1 < passes:
            rk = _resid(st, Val, d, r, transpose)
    return d


def _difference(adj, prop, u):
    """K'^-1 [(f' - f0) - (K' - K0) u] by refined solves, with K', f' the
    tangent and right-hand side assembled at prop and K0, f0 those of the
    last forward solve (kept), u its solution.

    Exactly, u(prop) - u = this + K'^-1 (f0 - K0 u): the base solve's
    residual is left out, so the derivative of u + this in prop is the
    tangent-linear derivative at the computed u, which is what the adjoint
    gradient differentiates.  The right-hand side is formed from
    differences that vanish outside the changed elements (the assembly is
    bitwise deterministic), without the O(eps |K| |u|) rounding of
    f' - K' u that a central difference would divide by the step."""
    st = adj.st
    if adj.Val0 is None or adj.f0 is None:
        raise fsils.SvfsiError("difference needs a forward solve first")
    st.prop = prop
    st.assemble()
    Val, f = st.Val.clone(), st.R.clone()
    r = (f - adj.f0) - _spmv(st, Val - adj.Val0, u)
    solve, _ = adj._solve_with(Val)
    return _refined_correction(st, Val, r, solve, adj.passes)


class LelasAdjoint:
    """Forward solve and adjoint gradient of LELAS with per-element modulus
    E (device tensor, nEl).  Keeps the unscaled tangent of the last forward
    solve for the transposed solve."""

    def __init__(self, prob: Problem, passes: int = 2):
        self.prob = prob
        self.st = NewtonStep(prob)
        self.Val0 = None
        self.f0 = None
        self.passes = passes  # residual-correction passes of every solve

    def _solve_with(self, Val):
        """rhs -> K^-1 rhs (FP64 device solve on a scaled copy of Val)."""
        st = self.st
        last = []

        def solve(rhs):
            st.Val.copy_(Val)
            st.R.copy_(rhs)
            ls = st.solve()
            if ls.RI.suc != 1:
                raise fsils.SvfsiError(f"solve did not converge ({ls.RI.itr} iterations)")
            last[:] = [ls]
            return st.R.clone()

        return solve, last

    def forward(self, E, keep: bool = True):
        """u = K(E)^-1 f (FP64 device solve refined by residual-correction
        passes); returns (u, ls of the first pass)."""
        st = self.st
        st.prop = E
        st.assemble()
        Val, f = st.Val.clone(), st.R.clone()
        if keep:
            self.Val0, self.f0 = Val, f
        solve, last = self._solve_with(Val)
        u = solve(f)
        ls = last[0]
        if self.passes > 1:
            u = u + _refined_correction(st, Val, _resid(st, Val, u, f), solve, self.passes - 1)
        return u, ls

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

    def correction(self, E, u):
        """d = K(E)^-1 (f - K(E) u), the change of the solution from u to
        u(E), by refined solves: the solver's relative tolerance applies to
        the (small) correction itself."""
        st = self.st
        st.prop = E
        st.assemble()
        Val, f = st.Val.clone(), st.R.clone()
        solve, _ = self._solve_with(Val)
        return _refined_correction(st, Val, _resid(st, Val, u, f), solve, self.passes)

    def solve_correction(self, E, u):
        """u(E) in correction form: u + K(E)^-1 (f - K(E) u)."""
        return u + self.correction(E, u)

    def difference(self, E, u):
        """K(E)^-1 [(f(E) - f0) - (K(E) - K0) u] against the last forward
        solve (u its solution): the change of u to first order in E - E0,
        without the base solve's residual."""
        return _difference(self, E, u)

    def grad_E(self, u, dLdu):
        """dL/dE_e for the last forward solve (u, its tangent in Val0); the
        transposed solve is refined like the forward one."""
        import torch

        st = self.st
        if self.Val0 is None:
            raise fsils.SvfsiError("grad_E needs a forward solve first")
        last = []

        def solve_t(rhs):
            lam = torch.empty_like(rhs)
            ls = st.make_ls()
            fsils.fsils_solve_t(st.lhs, ls, self.prob.dof, self.Val0, rhs.contiguous(), lam)
            if ls.RI.suc != 1:
                raise fsils.SvfsiError(f"adjoint solve did not converge ({ls.RI.itr} iterations)")
            last[:] = [ls] if not last else last
            return lam

        lam = solve_t(dLdu)
        for _ in range(self.passes - 1):
            lam = lam + solve_t(_resid(st, self.Val0, lam, dLdu, transpose=True))
        g = torch.empty(self.prob.mesh.nEl, dtype=torch.float64, device=u.device)
        st.dmesh.lelas_dkde_dot(self.prob.eq, lam, u, g)
        return g, lam, last[0]


class FluidViscosityAdjoint:
    """Linearised fluid step K(mu) du = R(mu) (FLUID3D VMS, per-element
    viscosity mu, device tensor nEl) and the adjoint gradient dL/dmu_e =
    lam_e^T (dR_e/dmu_e - dK_e/dmu_e du_e) with K^T lam = dL/du.  mixed
    selects the mixed-precision solver for the forward and transposed
    solves."""

    def __init__(self, prob: Problem, mixed: bool = False, passes: int = 2):
        self.prob = prob
        self.mixed = mixed
        self.st = NewtonStep(prob)
        self.Val0 = None
        self.f0 = None
        self.passes = passes  # residual-correction passes of every solve

    def _solve_with(self, Val):
        st = self.st
        last = []

        def solve(rhs):
            st.Val.copy_(Val)
            st.R.copy_(rhs)
            ls = st.solve(mixed=self.mixed)
            if ls.RI.suc != 1:
                raise fsils.SvfsiError(f"solve did not converge ({ls.RI.itr} iterations)")
            last[:] = [ls]
            return st.R.clone()

        return solve, last

    def forward(self, mu, keep: bool = True):
        """du = K(mu)^-1 R(mu), refined by residual-correction passes;
        returns (du, ls of the first pass)."""
        st = self.st
        st.prop = mu
        st.assemble()
        Val, f = st.Val.clone(), st.R.clone()
        if keep:
            self.Val0, self.f0 = Val, f
        solve, last = self._solve_with(Val)
        du = solve(f)
        ls = last[0]
        if self.passes > 1:
            du = du + _refined_correction(st, Val, _resid(st, Val, du, f), solve, self.passes - 1)
        return du, ls

    def correction(self, mu, du):
        """K(mu)^-1 (R(mu) - K(mu) du) by refined solves."""
        st = self.st
        st.prop = mu
        st.assemble()
        Val, f = st.Val.clone(), st.R.clone()
        solve, _ = self._solve_with(Val)
        return _refined_correction(st, Val, _resid(st, Val, du, f), solve, self.passes)

    def solve_correction(self, mu, du):
        """du(mu) in correction form: du + K(mu)^-1 (R(mu) - K(mu) du)."""
        return du + self.correction(mu, du)

    def difference(self, mu, du):
        """K(mu)^-1 [(R(mu) - R0) - (K(mu) - K0) du] against the last forward
        solve (du its solution): the change of du to first order in
        mu - mu0, without the base solve's residual."""
        return _difference(self, mu, du)

    def grad_mu(self, mu, du, dLdu):
        """dL/dmu_e for the last forward solve (its tangent kept in Val0); the
        transposed solve is refined like the forward one."""
        import torch

        st = self.st
        if self.Val0 is None:
            raise fsils.SvfsiError("grad_mu needs a forward solve first")
        last = []

        def solve_t(rhs):
            lam = torch.empty_like(rhs)
            ls = st.make_ls()
            fsils.fsils_solve_t(st.lhs, ls, self.prob.dof, self.Val0, rhs.contiguous(), lam, mixed=self.mixed)
            if ls.RI.suc != 1:
                raise fsils.SvfsiError(f"adjoint solve did not converge ({ls.RI.itr} iterations)")
            last[:] = [ls] if not last else last
            return lam

        lam = solve_t(dLdu)
        for _ in range(self.passes - 1):
            lam = lam + solve_t(_resid(st, self.Val0, lam, dLdu, transpose=True))
        g = torch.empty(self.prob.mesh.nEl, dtype=torch.float64, device=du.device)
        st.dmesh.fluid_dmu_dot(self.prob.eq, st.Ag, st.Yg, mu, lam, du, g)
        return g, lam, last[0]
