"""Pins of the oracle's BICGSV (svFSILS/BICGS.f:115-180): on the
diagonally scaled FLUID (nonsymmetric) and LELAS systems, FSILS_SOLVE with
LS_TYPE_BICGS converges to the dense solution of the same system, and its
reported final residual is the residual recomputed from the solution."""
import numpy as np
import pytest

from svfsi_amd import problems
from test_adjoint import dense


@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_bicgs_matches_dense_solution(orc, which):
    prob = problems.fluid_cylinder(h=0.6) if which == "fluid" else problems.lelas_cube(3)
    dof = prob.dof
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    st = dict(LS_type=795, relTol=1e-12, absTol=1e-30, mItr=5000)
    x, _, ls = orc.fsils_solve(rp, cp, dof, prob.dir_mask(), st, R, V)
    assert ls.suc == 1 and 0 < ls.itr < 5000
    Vs, Rs, W = orc.precond_diag(rp, cp, dof, V, R, prob.dir_mask())
    K = dense(Vs, rp, cp, prob.nNo, dof)
    free = np.abs(np.diag(K)) > 0
    y = np.linalg.solve(K + np.diag(~free * 1.0), Rs.ravel())
    xd = (W.ravel() * y).reshape(prob.nNo, dof)
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()
    assert ls.fNorm < 1e-12 * ls.iNorm
    # the updated residual against the true one, at 1e-8 (near 1e-12 the
    # recurrence drifts from the true residual by rounding, a known
    # property of BiCGStab, so the comparison is made well above it)
    x8, _, l8 = orc.fsils_solve(rp, cp, dof, prob.dir_mask(), dict(st, relTol=1e-8), R, V)
    yx = np.where(W != 0, x8 / np.where(W != 0, W, 1.0), 0.0)
    r = Rs - orc.spmv(rp, cp, dof, Vs, yx)
    assert l8.suc == 1 and np.linalg.norm(r) == pytest.approx(l8.fNorm, rel=1e-2)
