"""Pins of the oracle's Taylor-Hood (P2P1) pieces that the reference
fixtures do not isolate: GNNxx (svFSI/NN.f:1568-1678) and the TET10
second-derivative table (NN.f:1285-1295).

- On an affine tet10 the physical Hessian of every shape function is
  J^-T H_xi J^-1 (H_xi the constant reference Hessian, J = dx/dxi): GNNxx's
  6 x 6 solve must reproduce it.
- On a curved tet10 (a mid-edge node moved off its edge) the chain rule
  gains the x_xixi term: GNNxx must equal the derivative of GNN's physical
  gradient along x, taken by central differences in xi.
- The table: second differences of orc_shape_tet10's values.
The assembled Taylor-Hood system itself is pinned by the reference's own
output in tests/test_ref_golden.py (ref_fluid_th10_*)."""
import numpy as np
import pytest

from svfsi_amd import mesh

PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]  # xx, yy, zz, xy, yz, xz


def tet10_nodes(rng, curved=False):
    v = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]) + 0.15 * rng.normal(size=(4, 3))
    # the reference's node order: corners 1..4 with N1 = xi1 .. N4 = 1 - sum,
    # so corner a sits at xi = e_a (a = 1..3) and corner 4 at the origin
    x = np.zeros((10, 3))
    x[:4] = v
    for k, (a, b) in enumerate(mesh._MID):
        x[4 + k] = 0.5 * (x[a] + x[b])
    if curved:
        x[5] += np.array([0.08, -0.05, 0.06])
    return x


def hess_xi(orc, xi, h=1e-4):
    """Reference Hessian (10, 6) by second differences of the shape values."""
    out = np.zeros((10, 6))
    for c, (i, j) in enumerate(PAIRS):
        def N(d):
            return orc.shape_tet10(np.asarray(xi) + d)[0]
        ei, ej = np.eye(3)[i] * h, np.eye(3)[j] * h
        out[:, c] = (N(ei + ej) - N(ei - ej) - N(-ei + ej) + N(-ei - ej)) / (4 * h * h)
    return out


def gnn(orc, xi, xl):
    """(Nx (10, 3), xXi (3, 3)) at xi: GNN's physical gradient."""
    _, Nxi, _ = orc.shape_tet10(xi)
    J = xl.T @ Nxi            # J[k, i] = dx_k / dxi_i
    Nx = Nxi @ np.linalg.inv(J)
    return Nx, J


def test_tet10_second_derivative_table(orc):
    rng = np.random.default_rng(0)
    for xi in rng.uniform(0.05, 0.3, (4, 3)):
        _, _, Nxi2 = orc.shape_tet10(xi)
        assert np.allclose(Nxi2, hess_xi(orc, xi), atol=1e-6)


def test_gnnxx_affine_tet10_is_the_transformed_reference_hessian(orc):
    rng = np.random.default_rng(1)
    for _ in range(5):
        xl = tet10_nodes(rng)
        xi = rng.uniform(0.05, 0.3, 3)
        _, Nxi, Nxi2 = orc.shape_tet10(xi)
        Nx, J = gnn(orc, xi, xl)
        Nxx = orc.gnnxx(Nxi, Nxi2, xl, Nx)
        Ji = np.linalg.inv(J)
        for a in range(10):
            H = np.zeros((3, 3))
            for c, (i, j) in enumerate(PAIRS):
                H[i, j] = H[j, i] = Nxi2[a, c]
            Hx = Ji.T @ H @ Ji
            expect = [Hx[i, j] for i, j in PAIRS]
            assert np.allclose(Nxx[a], expect, rtol=1e-12, atol=1e-12 * np.abs(Hx).max())


def test_gnnxx_curved_tet10_matches_differences_of_the_gradient(orc):
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(5):
        xl = tet10_nodes(rng, curved=True)
        xi = rng.uniform(0.1, 0.25, 3)
        _, Nxi, Nxi2 = orc.shape_tet10(xi)
        Nx, J = gnn(orc, xi, xl)
        Nxx = orc.gnnxx(Nxi, Nxi2, xl, Nx)
        # d(Nx_i)/dxi_k by central differences, then along x: d/dx_j = sum_k dxi_k/dx_j d/dxi_k
        dNx = np.zeros((10, 3, 3))
        for k in range(3):
            e = np.eye(3)[k] * h
            dNx[:, :, k] = (gnn(orc, xi + e, xl)[0] - gnn(orc, xi - e, xl)[0]) / (2 * h)
        Ji = np.linalg.inv(J)  # Ji[k, j] = dxi_k / dx_j
        Hx = np.einsum("aik,kj->aij", dNx, Ji)
        for c, (i, j) in enumerate(PAIRS):
            assert np.allclose(Nxx[:, c], 0.5 * (Hx[:, i, j] + Hx[:, j, i]), rtol=1e-6,
                               atol=1e-6 * np.abs(Hx).max())
        # the curvature term matters here: the affine formula is off
        H0 = np.zeros((10, 3, 3))
        for c, (i, j) in enumerate(PAIRS):
            H0[:, i, j] = H0[:, j, i] = Nxi2[:, c]
        Haff = np.einsum("ki,akl,lj->aij", Ji, H0, Ji)
        assert np.abs(Haff - Hx).max() > 1e-3 * np.abs(Hx).max()
