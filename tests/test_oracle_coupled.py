"""Pins of the oracle's coupled Neumann faces (ADDBCMUL, svFSILS/ADDBCMUL.f:53-114;
BCPRE, GMRES.f:393-429 / NSSOLVER.f:307-338; valM, PRECOND.f:133-140):

- FSILS_SOLVE with GMRES and a resistance outlet converges to the dense
  solution of the scaled system plus the rank-1 term res vm vm^T on the
  outlet's velocity components (a dropped term, a wrong sign, a missing
  scaling of the face vector by W or a wrong node index fails it);
- the NS solver converges to the same dense solution;
- CGRADV ignores the coupled faces, as the reference does (bitwise the
  plain solve);
- PRE inverts ADD: with nS = vm.vm, (I + c_PRE v v^T)(I + res v v^T) = I
  (Sherman-Morrison), checked through the NS solver's one-iteration
  directions on a system whose momentum block is the identity."""
import numpy as np
import pytest

from svfsi_amd import problems
from test_adjoint import dense


def fluid_with_outlet(h=0.6, seed=3):
    prob = problems.fluid_cylinder(h=h)
    rng = np.random.default_rng(seed)
    prob.Yg[:, :3] += 0.1 * rng.normal(size=(prob.nNo, 3))
    ff, fe = prob.bflow_faces[0]
    nodes, val = problems.face_normal_integral(prob.mesh, ff, fe)
    return prob, nodes, val


def scaled(orc, prob):
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    Vs, Rs, W = orc.precond_diag(rp, cp, 4, V, R, prob.dir_mask())
    return rp, cp, V, R, Vs, Rs, W


def dense_coupled(Vs, Rs, W, rp, cp, n, faces):
    """Dense solve of (Ks + sum res u u^T) y = Rs, x = W y; u = val W on the
    face's nodes and velocity components."""
    K = dense(Vs, rp, cp, n, 4)
    for nodes, fd, val, res in faces:
        u = np.zeros((n, 4))
        u[nodes, :fd] = val * W[nodes, :fd]
        u = u.ravel()
        K = K + res * np.outer(u, u)
    free = np.abs(np.diag(K)) > 0
    K = K + np.diag(~free * 1.0)
    y = np.linalg.solve(K, Rs.ravel())
    return (W.ravel() * y).reshape(n, 4)


def res_for(nodes, val, W, target):
    """res with res * |vm|^2 = target (vm = val W)"""
    vm = val * W[nodes, :3]
    return target / float((vm * vm).sum())


@pytest.mark.parametrize("target", [1.0, 100.0])
def test_gmres_with_resistance_matches_dense(orc, target):
    prob, nodes, val = fluid_with_outlet()
    rp, cp, V, R, Vs, Rs, W = scaled(orc, prob)
    faces = [(nodes, 3, val, res_for(nodes, val, W, target))]
    st = dict(LS_type=797, relTol=1e-13, absTol=1e-30, mItr=400, sD=80)
    x, _, ls = orc.fsils_solve_cpl(rp, cp, 4, prob.dir_mask(), st, R, V, faces)
    assert ls.RI.suc == 1
    xd = dense_coupled(Vs, Rs, W, rp, cp, prob.nNo, faces)
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()
    # and it is not the uncoupled solution
    x0, _, _ = orc.fsils_solve(rp, cp, 4, prob.dir_mask(), st, R, V)
    assert np.abs(x - x0).max() > 1e-4 * np.abs(xd).max()


def test_ns_with_resistance_matches_dense(orc):
    prob, nodes, val = fluid_with_outlet()
    rp, cp, V, R, Vs, Rs, W = scaled(orc, prob)
    faces = [(nodes, 3, val, res_for(nodes, val, W, 10.0))]
    st = dict(LS_type=796, relTol=1e-10, absTol=1e-30, mItr=30, sD=50, GM=dict(relTol=1e-6, mItr=6),
              CG=dict(relTol=1e-6))
    x, _, ls = orc.fsils_solve_cpl(rp, cp, 4, prob.dir_mask(), st, R, V, faces)
    assert ls.RI.suc == 1
    xd = dense_coupled(Vs, Rs, W, rp, cp, prob.nNo, faces)
    assert np.abs(x - xd).max() <= 1e-6 * np.abs(xd).max()


def test_cg_ignores_coupled_faces(orc):
    prob = problems.lelas_cube(3)
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    ff, fe = prob.trac_faces[0][:2]
    nodes, val = problems.face_normal_integral(prob.mesh, ff, fe)
    faces = [(nodes, 3, val, 1e3)]
    st = dict(LS_type=798, relTol=1e-8, absTol=1e-30, mItr=500)
    x, _, ls = orc.fsils_solve_cpl(rp, cp, 3, prob.dir_mask(), st, R, V, faces)
    x0, _, l0 = orc.fsils_solve(rp, cp, 3, prob.dir_mask(), st, R, V)
    assert ls.RI.itr == l0.itr and np.array_equal(x, x0)


def test_rcs_with_coupled_faces_is_refused(orc):
    prob, nodes, val = fluid_with_outlet()
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    V, R = orc.assemble(prob, rp, cp)
    st = dict(LS_type=797, relTol=1e-8, absTol=1e-30, mItr=10, sD=20)
    with pytest.raises(RuntimeError, match="rc=-4"):
        orc.fsils_solve_cpl(rp, cp, 4, prob.dir_mask(), st, R, V, [(nodes, 3, val, 1.0)], prec=709)


def test_pre_inverts_add_through_ns(orc):
    """A block system whose scaled momentum block is the identity: K = I,
    G = D = 0 except a diagonal L.  With a resistance face, the momentum
    operator is I + res v v^T, and GMRES preconditioned by PRE
    (= its exact inverse when nS = v.v) finds U in one Krylov step: both
    GMRES calls of one outer iteration converge at their first iteration,
    and the solution is the dense one.  A wrong PRE coefficient (sign,
    missing 1 +, nS over the wrong components) needs more steps."""
    n = 30
    # a ring of nodes, each connected to itself and its neighbours
    rows = [sorted({i, (i - 1) % n, (i + 1) % n}) for i in range(n)]
    rp = np.zeros(n + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    cp = np.concatenate(rows).astype(np.int32)
    Val = np.zeros((len(cp), 16))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            if cp[j] == i:
                Val[j, [0, 5, 10]] = 1.0
                Val[j, 15] = 2.0
    rng = np.random.default_rng(0)
    R = rng.normal(size=(n, 4))
    nodes = np.arange(5, 12, dtype=np.int32)
    val = rng.normal(size=(len(nodes), 3))
    faces = [(nodes, 3, val, 7.0)]
    # GM relTol 1e-6: the Pythagorean h(2,1) = sqrt|1 - h(1,1)^2| of an exact
    # first step is ~1e-8 from rounding, not 0
    st = dict(LS_type=796, relTol=1e-12, absTol=1e-30, mItr=1, sD=20, GM=dict(relTol=1e-6, mItr=1),
              CG=dict(relTol=1e-12))
    x, _, ls = orc.fsils_solve_cpl(rp, cp, 4, None, st, R, Val, faces)
    # diagonal scaling leaves K = I (W = 1 on the momentum rows)
    # GMRES is called twice per outer iteration; each converges in one step
    assert ls.GM.itr == 2, ls.GM.itr
    u = np.zeros((n, 3))
    u[nodes] = val
    A = np.eye(3 * n) + 7.0 * np.outer(u.ravel(), u.ravel())
    U = np.linalg.solve(A, R[:, :3].ravel()).reshape(n, 3)
    W = 1.0 / np.sqrt(np.array([1.0, 1.0, 1.0, 2.0]))
    assert np.allclose(x[:, :3], W[:3] * U, rtol=1e-10, atol=1e-12)
