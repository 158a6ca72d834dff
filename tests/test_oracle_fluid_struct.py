"""Pins the oracle's FLUID3D_M/C and STRUCT3D + GETPK2CC by finite
differences of its own residual and by closed forms.

The unknown of a Newton step is the acceleration increment: Ag moves by
am*dA, Yg by af*gam*dt*dA, Dg by af*beta*dt^2*dA (PICI, svFSI/PIC.f:153-157),
so the tangent is K = am dR/dAg + af*gam*dt dR/dYg (fluid) or
am dR/dAg + af*beta*dt^2 dR/dDg (solids).

FLUID: the VMS tangent of FLUID.f is not the exact Jacobian in general (the
dependence of tauM, tauC, tauB on u and the (du.grad)u term are dropped).
At a uniform velocity U with Ag = 0 and constant pressure, the fine-scale
velocity up vanishes and every dropped term multiplies a vanishing factor,
so there the tangent must equal the Jacobian: that pins the mass, viscous,
convective, SUPG/PSPG and LSIC blocks and the tauM / tauC / metric values.
The tauB terms and the residual at general states are not pinned by these
checks (parity unpinned for them).

STRUCT: the neo-Hookean tangent is the consistent linearisation (solid
viscosity 0), so it must match the finite difference everywhere."""
import numpy as np
import pytest

from svfsi_amd import mesh, problems
from svfsi_amd._lib import SVFSI_PHYS_FLUID, SVFSI_PHYS_STRUCT


def dense(rp, cp, dof, Val):
    nNo = len(rp) - 1
    K = np.zeros((nNo * dof, nNo * dof))
    for r in range(nNo):
        for k in range(rp[r], rp[r + 1]):
            c = cp[k]
            K[r * dof:(r + 1) * dof, c * dof:(c + 1) * dof] = Val[k].reshape(dof, dof)
    return K


def fluid_problem(U, p0=0.3, dt=0.05):
    m = mesh.cube_tet4(2)
    rng = np.random.default_rng(0)
    m.x = m.x + rng.uniform(-0.05, 0.05, m.x.shape) * (m.x * (1 - m.x) > 0)  # jitter interior nodes
    eq = dict(problems.gen_alpha(0.2, False), dt=dt, rho=1.06, mu=0.04, bstab=0.2)
    Y = np.zeros((m.nNo, 4))
    Y[:, :3] = U
    Y[:, 3] = p0
    return problems.Problem(name="fd", phys=SVFSI_PHYS_FLUID, dof=4, mesh=m, eq=eq, Ag=np.zeros((m.nNo, 4)),
                            Yg=Y, Dg=np.zeros((m.nNo, 4)), dir_faces=[])


@pytest.mark.parametrize("U", [(0.0, 0.0, 0.0), (1.0, -0.5, 0.25), (3.0, 2.0, -1.0)])
def test_fluid_tangent_equals_fd_at_uniform_flow(orc, U):
    prob = fluid_problem(np.array(U))
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    Val, R = orc.assemble(prob, rp, cp)
    K = dense(rp, cp, 4, Val)
    eq = prob.eq
    T1 = eq["af"] * eq["gam"] * eq["dt"]
    rng = np.random.default_rng(1)
    cols = rng.choice(prob.nNo * 4, size=24, replace=False)
    h = 1e-6
    for col in cols:
        b, c = divmod(col, 4)
        Rs = []
        for sgn in (1, -1):
            p2 = fluid_problem(np.array(U))
            p2.mesh = prob.mesh
            p2.Ag = prob.Ag.copy()
            p2.Yg = prob.Yg.copy()
            p2.Ag[b, c] += sgn * eq["am"] * h
            p2.Yg[b, c] += sgn * T1 * h
            Rs.append(orc.assemble(p2, rp, cp)[1].ravel())
        fd = (Rs[0] - Rs[1]) / (2 * h)
        assert np.allclose(K[:, col], fd, rtol=2e-5, atol=2e-6 * np.abs(K).max()), (U, b, c)


def test_fluid_rest_state_residual(orc):
    # u = 0, a = 0, constant p: only boundary nodes feel the pressure
    prob = fluid_problem(np.zeros(3), p0=2.0)
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    _, R = orc.assemble(prob, rp, cp)
    x = prob.mesh.x
    interior = ((x > 0) & (x < 1)).all(axis=1)
    assert np.abs(R[interior]).max() < 1e-13
    assert np.abs(R[:, 3]).max() < 1e-13
    # the boundary terms add up to -p * (closed-surface integral of n) = 0
    assert np.abs(R[:, :3].sum(axis=0)).max() < 1e-12


def struct_problem(n=1, quad=1, seed=0, amp=0.05):
    prob = problems.struct_cube(n=n, quad=quad, seed=seed, amp=amp, dt=0.1)
    prob.Ag = np.random.default_rng(seed + 7).normal(size=prob.Ag.shape)
    return prob


@pytest.mark.parametrize("quad", [0, 1])
def test_struct_tangent_is_consistent(orc, quad):
    prob = struct_problem(quad=quad)
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    Val, R = orc.assemble(prob, rp, cp)
    K = dense(rp, cp, 3, Val)
    eq = prob.eq
    afu = eq["af"] * eq["beta"] * eq["dt"] ** 2
    rng = np.random.default_rng(3)
    cols = rng.choice(prob.nNo * 3, size=20, replace=False)
    h = 1e-5
    for col in cols:
        b, c = divmod(col, 3)
        Rs = []
        for sgn in (1, -1):
            p2 = struct_problem(quad=quad)
            p2.Ag[b, c] += sgn * eq["am"] * h
            p2.Dg[b, c] += sgn * afu * h
            Rs.append(orc.assemble(p2, rp, cp)[1].ravel())
        fd = (Rs[0] - Rs[1]) / (2 * h)
        assert np.allclose(K[:, col], fd, rtol=1e-5, atol=1e-7 * np.abs(K).max()), (quad, b, c)


def _pk2(orc, eq, F):
    import ctypes as C

    S = np.zeros((3, 3))
    Dm = np.zeros((6, 6))
    e = orc.eq_struct(eq)
    orc.lib().orc_getpk2cc_nhk(C.byref(e), orc.P(np.ascontiguousarray(F, dtype=float)), orc.P(S), orc.P(Dm))
    return S, Dm


def test_getpk2cc_reference_state(orc):
    C10, K = 1e5, 1e7
    mu = 2 * C10
    for vol in (1, 2, 3):
        S, Dm = _pk2(orc, dict(C10=C10, Kpen=K, volType=vol), np.eye(3))
        assert np.abs(S).max() < 1e-9
        exp = np.zeros((6, 6))
        exp[:3, :3] = K - 2 * mu / 3
        exp[:3, :3] += 2 * mu * np.eye(3)
        exp[3:, 3:] = mu * np.eye(3)
        # Quad: pl = K(2J-1) = K at J = 1; ST91: pl = K J = K; M94: pl = K
        assert np.allclose(Dm, exp, rtol=1e-12, atol=1e-6), vol


def test_getpk2cc_objective_and_symmetric(orc):
    rng = np.random.default_rng(2)
    F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    eq = dict(C10=1e5, Kpen=1e7, volType=2)
    S1, D1 = _pk2(orc, eq, F)
    S2, D2 = _pk2(orc, eq, Q @ F)
    assert np.allclose(S1, S2, rtol=1e-10, atol=1e-6)
    assert np.allclose(D1, D2, rtol=1e-10, atol=1e-3)
    assert np.allclose(S1, S1.T, atol=1e-8)
    assert np.allclose(D1, D1.T, atol=1e-6)
