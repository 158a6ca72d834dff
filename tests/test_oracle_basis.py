"""Pins the oracle's quadrature, shape functions and GNN (svFSI/NN.f) with
closed forms: exact integrals of monomials over the reference tet,
Kronecker-delta property at the nodes, and affine reproduction."""
import ctypes as C
from math import factorial

import numpy as np
import pytest


def _rule(orc, name, nG):
    w = np.zeros(nG)
    xi = np.zeros((nG, 3))
    getattr(orc.lib(), name)(orc.P(w), orc.P(xi))
    return w, xi


def _exact(a, b, c):
    # int over {xi >= 0, sum xi <= 1} of x^a y^b z^c
    return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)


@pytest.mark.parametrize("name,nG,deg,tol", [("orc_rule_tet4", 4, 2, 1e-15), ("orc_rule_tet10", 15, 5, 5e-14)])
def test_tet_rules_integrate_polynomials(orc, name, nG, deg, tol):
    w, xi = _rule(orc, name, nG)
    assert abs(w.sum() - 1.0 / 6.0) < tol
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                q = np.sum(w * xi[:, 0] ** a * xi[:, 1] ** b * xi[:, 2] ** c)
                assert abs(q - _exact(a, b, c)) < tol, (a, b, c)


def test_tri3_rule(orc):
    w = np.zeros(3)
    xi = np.zeros((3, 2))
    orc.lib().orc_rule_tri3(orc.P(w), orc.P(xi))
    assert abs(w.sum() - 0.5) < 1e-16
    # exact for quadratics: int x^2 = 1/12, int xy = 1/24 over the unit triangle
    assert abs(np.sum(w * xi[:, 0] ** 2) - 1 / 12) < 1e-15
    assert abs(np.sum(w * xi[:, 0] * xi[:, 1]) - 1 / 24) < 1e-15


TET10_NODES = np.array([
    [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0],
    [.5, .5, 0], [0, .5, .5], [.5, 0, .5], [.5, 0, 0], [0, .5, 0], [0, 0, .5]])


def _shape(orc, name, n, xi):
    N = np.zeros(n)
    Nxi = np.zeros((n, 3))
    getattr(orc.lib(), name)(orc.P(np.ascontiguousarray(xi, dtype=float)), orc.P(N), orc.P(Nxi))
    return N, Nxi


def test_tet10_kronecker_and_partition(orc):
    # node order: corners xi1, xi2, xi3, 1-sum; mid-edges (1,2),(2,3),(1,3),(1,4),(2,4),(3,4)
    for b in range(10):
        N, _ = _shape(orc, "orc_shape_tet10", 10, TET10_NODES[b])
        assert np.allclose(N, np.eye(10)[b], atol=1e-15), b
    rng = np.random.default_rng(0)
    for _ in range(5):
        x = rng.dirichlet(np.ones(4))[:3]
        N, Nxi = _shape(orc, "orc_shape_tet10", 10, x)
        assert abs(N.sum() - 1) < 1e-14
        assert np.allclose(Nxi.sum(0), 0, atol=1e-13)
        # derivative check by central differences
        for j in range(3):
            e = np.zeros(3)
            e[j] = 1e-6
            Np, _ = _shape(orc, "orc_shape_tet10", 10, x + e)
            Nm, _ = _shape(orc, "orc_shape_tet10", 10, x - e)
            assert np.allclose((Np - Nm) / 2e-6, Nxi[:, j], atol=1e-8)


def test_tet4_shape(orc):
    for b, node in enumerate(TET10_NODES[:4]):
        N, Nxi = _shape(orc, "orc_shape_tet4", 4, node)
        assert np.allclose(N, np.eye(4)[b])
    assert np.allclose(Nxi.sum(0), 0)


def _gnn(orc, Nxi, xl):
    n = len(xl)
    Nx = np.zeros((n, 3))
    J = C.c_double()
    ks = np.zeros((3, 3))
    orc.lib().orc_gnn(n, orc.P(np.ascontiguousarray(Nxi)), orc.P(np.ascontiguousarray(xl)), orc.P(Nx), C.byref(J), orc.P(ks))
    return Nx, J.value, ks


def test_gnn_affine(orc):
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3)) + 3 * np.eye(3)
    b = rng.normal(size=3)
    # reference nodes of tet4 in xi: node1 = e1, node2 = e2, node3 = e3, node4 = 0
    ref = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0.0]])
    xl = ref @ A.T + b
    _, Nxi = _shape(orc, "orc_shape_tet4", 4, np.array([0.2, 0.3, 0.1]))
    Nx, J, ks = _gnn(orc, Nxi, xl)
    assert abs(J - np.linalg.det(A)) < 1e-12 * abs(J)
    # physical gradients reproduce linear fields: sum_a Nx[a,j] x_a,i = delta_ij
    assert np.allclose(Nx.T @ xl, np.eye(3), atol=1e-12)
    # metric ks = (dxi/dx)^T (dxi/dx) = (A^-1)^T A^-1
    Ai = np.linalg.inv(A)
    assert np.allclose(ks, Ai.T @ Ai, rtol=1e-12, atol=1e-12)
    # tet10 on the same affine map: identical Jacobian at every point
    _, Nxi10 = _shape(orc, "orc_shape_tet10", 10, np.array([0.1, 0.2, 0.3]))
    xl10 = TET10_NODES @ A.T + b
    Nx10, J10, _ = _gnn(orc, Nxi10, xl10)
    assert abs(J10 - J) < 1e-12 * abs(J)
    assert np.allclose(Nx10.T @ xl10, np.eye(3), atol=1e-12)


def test_iszero(orc):
    f = orc.lib().orc_iszero
    f.argtypes = [C.c_double]
    eps = np.finfo(float).eps
    assert f(0.0) == 1
    assert f(1e-40) == 1
    assert f(eps) == 0
    assert f(1.0) == 0
