"""Pins the oracle's LELAS3D / CONSTRUCT_LELAS (svFSI/LELAS.f:38-242) with
the textbook linear-tet element: K_e = T1 V B^T D B + am rho M (x) I3 with
the consistent mass M_ab = V (1 + delta_ab) / 20, internal force = K_stiff d,
inertial force = rho M a, and rigid-body modes in the stiffness null space.
Also pins the traction face term (SETBCTRACL) against the closed form
lR_a = -h area / 3 for a uniform traction."""
import ctypes as C

import numpy as np
import pytest

from svfsi_amd.problems import gen_alpha


def one_tet(rng):
    A = rng.normal(size=(3, 3)) + 2.5 * np.eye(3)
    ref = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0.0]])
    x = ref @ A.T + rng.normal(size=3)
    if np.linalg.det(A) < 0:
        x[[0, 1]] = x[[1, 0]]
    return x, np.array([[0, 1, 2, 3]], np.int32)


def element(orc, x, ien, eq, a=None, d=None, E=None):
    rp, cp = orc.lhsa(4, ien)
    Val = np.zeros((len(cp), 9))
    R = np.zeros((4, 3))
    A = np.zeros((4, 3)) if a is None else a
    D = np.zeros((4, 3)) if d is None else d
    e = orc.eq_struct(eq)
    Ee = None if E is None else orc.P(np.array([E], float))
    orc.lib().orc_construct_lelas(4, 1, orc.P(ien), orc.P(np.ascontiguousarray(x)), orc.P(orc.f64(A)),
                                  orc.P(orc.f64(D)), 3, C.byref(e), Ee,
                                  orc.P(rp), orc.P(cp), orc.P(Val), orc.P(R))
    # dense 12 x 12 from the block pattern (1 element -> full 4x4 blocks)
    K = np.zeros((12, 12))
    for r in range(4):
        for k in range(rp[r], rp[r + 1]):
            c = cp[k]
            K[3 * r:3 * r + 3, 3 * c:3 * c + 3] = Val[k].reshape(3, 3)
    return K, R.ravel()


def textbook(x, E, nu):
    # B for linear tet: grad N from the inverse of [1 x y z]
    Cm = np.hstack([np.ones((4, 1)), x])
    V = abs(np.linalg.det(Cm)) / 6.0
    G = np.linalg.inv(Cm)[1:, :]  # 3 x 4: dN_a/dx_j
    B = np.zeros((6, 12))
    for a in range(4):
        gx, gy, gz = G[:, a]
        B[0, 3 * a] = gx
        B[1, 3 * a + 1] = gy
        B[2, 3 * a + 2] = gz
        B[3, 3 * a], B[3, 3 * a + 1] = gy, gx        # gamma_xy
        B[4, 3 * a + 1], B[4, 3 * a + 2] = gz, gy    # gamma_yz
        B[5, 3 * a], B[5, 3 * a + 2] = gz, gx        # gamma_zx
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[:3, :3] += 2 * mu * np.eye(3)
    D[3:, 3:] = mu * np.eye(3)
    M = V * (np.ones((4, 4)) + np.eye(4)) / 20.0
    return V * B.T @ D @ B, np.kron(M, np.eye(3)), V


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lelas_tangent_closed_form(orc, seed):
    rng = np.random.default_rng(seed)
    x, ien = one_tet(rng)
    E, nu, rho, dt = 2.0e6, 0.3, 1.7, 0.5
    eq = dict(gen_alpha(0.2, True), dt=dt, rho=rho, E=E, nu=nu)
    T1 = eq["af"] * eq["beta"] * dt * dt
    K, _ = element(orc, x, ien, eq)
    Ks, M, V = textbook(x, E, nu)
    expect = T1 * Ks + eq["am"] * rho * M
    assert np.allclose(K, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())
    assert np.allclose(K, K.T, rtol=1e-12, atol=1e-12 * np.abs(K).max())


def test_lelas_residual_closed_form(orc):
    rng = np.random.default_rng(5)
    x, ien = one_tet(rng)
    E, nu, rho = 3.0e5, 0.25, 2.0
    eq = dict(gen_alpha(0.2, True), dt=1.0, rho=rho, E=E, nu=nu, fb=(0.3, -0.2, 0.1))
    Ks, M, V = textbook(x, E, nu)
    d = rng.normal(size=(4, 3))
    a = rng.normal(size=(4, 3))
    _, R = element(orc, x, ien, eq, a=a, d=d)
    f = np.tile(eq["fb"], 4)
    # lR = rho*M*(a - f) + K_stiff d   (LELAS.f:161-215; ud = -f + sum N a)
    expect = rho * M @ (a.ravel() - f) + Ks @ d.ravel()
    assert np.allclose(R, expect, rtol=1e-11, atol=1e-11 * np.abs(expect).max())


def test_lelas_rigid_body_modes(orc):
    rng = np.random.default_rng(7)
    x, ien = one_tet(rng)
    eq = dict(gen_alpha(0.2, True), dt=1.0, rho=0.0, E=1e6, nu=0.3)
    K, _ = element(orc, x, ien, eq)
    modes = []
    for j in range(3):
        t = np.zeros((4, 3))
        t[:, j] = 1
        modes.append(t.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, x).ravel())
    for r in modes:
        assert np.abs(K @ r).max() < 1e-9 * np.abs(K).max()


def test_lelas_per_element_modulus_scales(orc):
    rng = np.random.default_rng(8)
    x, ien = one_tet(rng)
    eq = dict(gen_alpha(0.2, True), dt=1.0, rho=0.0, E=1.0, nu=0.3)
    K1, _ = element(orc, x, ien, eq, E=1.0)
    K5, _ = element(orc, x, ien, eq, E=5.0)
    assert np.allclose(K5, 5 * K1, rtol=1e-13)


def test_traction_closed_form(orc):
    from svfsi_amd import mesh

    m = mesh.cube_tet4(2)
    rp, cp = orc.lhsa(m.nNo, m.ien)
    f, e = mesh.boundary_faces(m.ien)
    sel = (m.x[f][:, :, 0] == 1.0).all(1)
    f, e = f[sel], e[sel]
    h = np.array([1e3, -2.0, 5.0])
    Val = np.zeros((len(cp), 9))
    R = np.zeros((m.nNo, 3))
    orc.lib().orc_traction_tri3(len(f), orc.P(orc.i32(f)), orc.P(orc.i32(e)), orc.P(orc.i32(m.ien)), 4,
                                orc.P(orc.f64(m.x)), orc.P(h), 3, orc.P(rp), orc.P(cp), orc.P(Val), orc.P(R))
    expect = np.zeros((m.nNo, 3))
    for tri in f:
        p = m.x[tri]
        area = 0.5 * np.linalg.norm(np.cross(p[0] - p[2], p[1] - p[2]))
        for a in tri:
            expect[a] -= h * area / 3.0
    assert np.allclose(R, expect, rtol=1e-14, atol=1e-12)
    assert abs(R[:, 0].sum() + 1e3) < 1e-9  # total load = -h * area(x = 1) = -h
    assert not Val.any()
