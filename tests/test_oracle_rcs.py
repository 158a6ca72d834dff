"""Pins of the oracle's PRECONDRCS (svFSILS/PRECOND.f:150-368).

- Closed form on a diagonal system: one sweep scales every |d| to 1, so
  W1 = W2 = |d|^(-1/2), the scaled entries are sign(d), and the next sweep
  sees unit rows and stops.
- The sweep count follows the reference's loop: a sweep whose max |1 - W|
  is not below 2 is followed by another one, at most 10 in all.
- Dirichlet dofs: rows and columns removed, diagonal one, right-hand side
  zero.
- Invariance: x = W2 y, y the solution of the scaled system, solves the
  original system; FSILS_SOLVE with RCS reaches the same x as with the
  diagonal preconditioner (LELAS with CG, FLUID with GMRES)."""
import numpy as np
import pytest

from svfsi_amd import problems


def diag_system(d):
    nNo, dof = d.shape
    rp = np.arange(nNo + 1, dtype=np.int32)
    cp = np.arange(nNo, dtype=np.int32)
    V = np.zeros((nNo, dof * dof))
    for i in range(dof):
        V[:, i * dof + i] = d[:, i]
    return rp, cp, V


def test_rcs_closed_form_on_diagonal(orc):
    rng = np.random.default_rng(3)
    d = rng.uniform(0.5, 2.5, (7, 3)) * rng.choice([-1.0, 1.0], (7, 3))
    rp, cp, V = diag_system(d)
    R = rng.normal(size=(7, 3))
    Vs, Rs, W1, W2 = orc.precond_rcs(rp, cp, 3, V, R, None)
    s = np.abs(d) ** -0.5
    assert np.allclose(W1, s, rtol=1e-15) and np.allclose(W2, s, rtol=1e-15)
    for i in range(3):
        assert np.allclose(Vs[:, i * 3 + i], np.sign(d[:, i]), rtol=1e-15)
    assert np.allclose(Rs, s * R, rtol=1e-15)


def test_rcs_sweeps_until_rows_are_balanced(orc):
    # |d| up to 1e6: sweep 1 takes it to 1 (max |1 - W| = 1e6 - 1 >= 2), so
    # sweep 2 runs with unit rows (W = 1) and changes nothing more
    d = np.array([[1e6, 4.0], [9.0, 1e-4]])
    rp, cp, V = diag_system(d)
    Vs, Rs, W1, W2 = orc.precond_rcs(rp, cp, 2, V, np.ones((2, 2)), None)
    assert np.allclose(W1, np.abs(d) ** -0.5, rtol=1e-15)
    assert np.allclose(Vs[:, [0, 3]], 1.0, rtol=1e-15)


def test_rcs_dirichlet_structure(orc):
    prob = problems.lelas_cube(3)
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    mask = prob.dir_mask()
    Vs, Rs, W1, W2 = orc.precond_rcs(rp, cp, 3, V, R, mask)
    nodes = np.where(mask[:, 0] == 0.0)[0]
    assert len(nodes) > 0
    assert np.all(Rs[nodes] == 0.0)
    for a in nodes[:5]:
        for k in range(rp[a], rp[a + 1]):
            blk = Vs[k].reshape(3, 3)
            if cp[k] == a:
                assert np.allclose(blk, np.eye(3) * blk[0, 0]) and blk[0, 0] > 0
            else:
                assert np.all(blk == 0.0)
    # equilibrated: every scalar row's largest entry is within the stop band
    rowmax = np.zeros((prob.nNo, 3))
    for a in range(prob.nNo):
        for k in range(rp[a], rp[a + 1]):
            rowmax[a] = np.maximum(rowmax[a], np.abs(Vs[k].reshape(3, 3)).max(axis=1))
    assert rowmax.max() < 3.0


@pytest.mark.parametrize("which", ["lelas", "fluid"])
def test_rcs_solve_matches_diagonal_preconditioner(orc, which):
    prob = problems.lelas_cube(4) if which == "lelas" else problems.fluid_cylinder(h=0.6)
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    ls = dict(prob.ls, relTol=1e-10, absTol=1e-30, mItr=5000)
    x1, _, s1 = orc.fsils_solve(rp, cp, prob.dof, prob.dir_mask(), ls, R, V, prec=701)
    x2, _, s2 = orc.fsils_solve(rp, cp, prob.dof, prob.dir_mask(), ls, R, V, prec=709)
    assert s1.suc == 1 and s2.suc == 1
    assert np.abs(x2 - x1).max() <= 1e-6 * np.abs(x1).max()
    # x = W2 y solves the original (masked) system
    free = prob.dir_mask() != 0.0
    r = (R - orc.spmv(rp, cp, prob.dof, V, x2))[free]
    assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(R[free])
