"""Pins the oracle's svFSILS restatement: SPARMULVV against dense products,
PRECONDDIAG against its closed form, CGRADV/GMRESV against dense solves and
against exact-arithmetic iteration counts (k distinct eigenvalues -> k
iterations), FSILS_SOLVE's Dirichlet masking against the reduced system, and
GMRESV's absTol early return."""
import numpy as np
import pytest

from svfsi_amd import mesh


def pattern(n=3):
    m = mesh.cube_tet4(n)
    return m


def dense(rp, cp, dof, Val):
    nNo = len(rp) - 1
    K = np.zeros((nNo * dof, nNo * dof))
    for r in range(nNo):
        for k in range(rp[r], rp[r + 1]):
            c = cp[k]
            K[r * dof:(r + 1) * dof, c * dof:(c + 1) * dof] = Val[k].reshape(dof, dof)
    return K


def blocks_from_dense(rp, cp, dof, K):
    Val = np.zeros((len(cp), dof * dof))
    for r in range(len(rp) - 1):
        for k in range(rp[r], rp[r + 1]):
            c = cp[k]
            Val[k] = K[r * dof:(r + 1) * dof, c * dof:(c + 1) * dof].ravel()
    return Val


@pytest.mark.parametrize("dof", [2, 3, 4])
def test_spmv_dense(orc, dof):
    m = pattern(3)
    rp, cp = orc.lhsa(m.nNo, m.ien)
    rng = np.random.default_rng(dof)
    Val = rng.normal(size=(len(cp), dof * dof))
    U = rng.normal(size=(m.nNo, dof))
    K = dense(rp, cp, dof, Val)
    assert np.allclose(orc.spmv(rp, cp, dof, Val, U).ravel(), K @ U.ravel(), rtol=1e-13, atol=1e-12)
    assert np.allclose(orc.spmv_t(rp, cp, dof, Val, U).ravel(), K.T @ U.ravel(), rtol=1e-13, atol=1e-12)


def test_precond_closed_form(orc):
    m = pattern(2)
    dof = 3
    rp, cp = orc.lhsa(m.nNo, m.ien)
    rng = np.random.default_rng(3)
    Val = rng.normal(size=(len(cp), dof * dof))
    R = rng.normal(size=(m.nNo, dof))
    mask = np.ones((m.nNo, dof))
    mask[:4, :2] = 0.0
    K = dense(rp, cp, dof, Val)
    # one zero diagonal entry -> W = 1 there (PRECOND.f:104-109)
    k0 = [k for k in range(rp[5], rp[6]) if cp[k] == 5][0]
    Val[k0, 0] = 0.0
    K = dense(rp, cp, dof, Val)
    d = np.diag(K).copy()
    d[d == 0] = 1.0
    W = mask.ravel() / np.sqrt(np.abs(d))
    V2, R2, W2 = orc.precond_diag(rp, cp, dof, Val, R, mask)
    assert np.allclose(W2.ravel(), W, rtol=1e-15)
    assert np.allclose(dense(rp, cp, dof, V2), np.diag(W) @ K @ np.diag(W), rtol=1e-14, atol=1e-14)
    assert np.allclose(R2.ravel(), W * R.ravel(), rtol=1e-15)


def spd_blocks(orc, dof, eigs, seed):
    """Block-diagonal SPD matrix on a mesh pattern whose spectrum is `eigs`."""
    m = pattern(2)
    rp, cp = orc.lhsa(m.nNo, m.ien)
    n = m.nNo * dof
    rng = np.random.default_rng(seed)
    lam = rng.choice(eigs, size=n)
    K = np.diag(lam)
    return rp, cp, K, rng


def test_cg_exact_iterations_distinct_eigs(orc):
    dof = 3
    rp, cp, K, rng = spd_blocks(orc, dof, np.array([1.0, 2.0, 5.0]), 0)
    Val = blocks_from_dense(rp, cp, dof, K)
    R = rng.normal(size=(len(rp) - 1, dof))
    x, s = orc.cgradv(rp, cp, dof, Val, R, dict(relTol=1e-10, absTol=1e-30, mItr=100))
    assert s.suc == 1
    assert s.itr == 3
    assert np.allclose(K @ x.ravel(), R.ravel(), atol=1e-9)


def test_gmres_exact_iterations_distinct_eigs(orc):
    dof = 2
    rp, cp, K, rng = spd_blocks(orc, dof, np.array([1.0, 3.0, 4.0, 9.0]), 1)
    Val = blocks_from_dense(rp, cp, dof, K)
    R = rng.normal(size=(len(rp) - 1, dof))
    # relTol 1e-6: after the 4th step the Krylov space is exhausted and the
    # Pythagorean h(5,4) = sqrt(|<w,w> - sum h^2|) (GMRES.f:345-347) is
    # rounding noise of size ~sqrt(eps)|w|, so the estimate drops by ~1e-8,
    # not to zero
    x, s = orc.gmresv(rp, cp, dof, Val, R, dict(relTol=1e-6, absTol=1e-30, mItr=10, sD=30))
    assert s.suc == 1
    # one restart cycle: 1 (cycle start) + 4 Arnoldi steps
    assert s.itr == 1 + 4
    assert np.allclose(K @ x.ravel(), R.ravel(), atol=1e-6)


def test_cg_and_gmres_dense_solve(orc):
    m = pattern(3)
    dof = 3
    rp, cp = orc.lhsa(m.nNo, m.ien)
    rng = np.random.default_rng(4)
    Val = rng.normal(size=(len(cp), dof * dof)) * 0.1
    K = dense(rp, cp, dof, Val)
    n = K.shape[0]
    Ks = 0.5 * (K + K.T)
    Ks += (1.0 - np.linalg.eigvalsh(Ks).min()) * np.eye(n)  # SPD, same pattern
    Kn = K + 4.0 * np.eye(n)
    R = rng.normal(size=(m.nNo, dof))
    xs, s = orc.cgradv(rp, cp, dof, blocks_from_dense(rp, cp, dof, Ks), R,
                       dict(relTol=1e-12, absTol=1e-30, mItr=5000))
    assert s.suc == 1
    assert np.allclose(xs.ravel(), np.linalg.solve(Ks, R.ravel()), rtol=1e-9, atol=1e-10)
    # GMRES with restarts (sD = 7): the restart logic must still converge
    xg, s = orc.gmresv(rp, cp, dof, blocks_from_dense(rp, cp, dof, Kn), R,
                       dict(relTol=1e-12, absTol=1e-30, mItr=500, sD=7))
    assert s.suc == 1
    assert np.allclose(xg.ravel(), np.linalg.solve(Kn, R.ravel()), rtol=1e-9, atol=1e-10)
    assert s.itr > 8  # it did restart


def test_fsils_solve_dirichlet_reduced_system(orc):
    m = pattern(3)
    dof = 3
    rp, cp = orc.lhsa(m.nNo, m.ien)
    rng = np.random.default_rng(6)
    Val = rng.normal(size=(len(cp), dof * dof)) * 0.1
    K = dense(rp, cp, dof, Val)
    n = K.shape[0]
    K = K + 3.0 * np.eye(n)
    Val = blocks_from_dense(rp, cp, dof, K)
    R = rng.normal(size=(m.nNo, dof))
    mask = np.ones((m.nNo, dof))
    fixed = m.x[:, 0] == 0.0
    mask[fixed] = 0.0
    free = mask.ravel() > 0
    for LS in (797, 798):
        KK = K if LS == 797 else 0.5 * (K + K.T) + 3.0 * np.eye(n)
        VV = blocks_from_dense(rp, cp, dof, KK)
        x, Vs, s = orc.fsils_solve(rp, cp, dof, mask, dict(LS_type=LS, relTol=1e-12, absTol=1e-30, mItr=3000, sD=60), R, VV)
        assert s.suc == 1
        assert np.all(x.ravel()[~free] == 0.0)
        xf = np.linalg.solve(KK[np.ix_(free, free)], R.ravel()[free])
        # GMRESV's Pythagorean h(i+1,i) limits the attainable accuracy to ~sqrt(eps)
        assert np.allclose(x.ravel()[free], xf, rtol=1e-6, atol=1e-7)


def test_gmres_abstol_early_return_keeps_rhs(orc):
    m = pattern(2)
    dof = 2
    rp, cp = orc.lhsa(m.nNo, m.ien)
    Val = np.tile(np.eye(dof).ravel(), (len(cp), 1))
    R = np.full((m.nNo, dof), 1e-14)
    x, s = orc.gmresv(rp, cp, dof, Val, R, dict(relTol=0.1, absTol=1e-10, mItr=4, sD=10))
    assert s.itr == 0 and s.suc == 0
    assert np.array_equal(x, R)  # GMRES.f:307-311 returns before R = X
