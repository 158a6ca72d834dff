"""Adjoint gradient dL/dz of L = 0.5 ||u - u_obs||^2, u = K(E)^-1 f,
E_e = E0 exp(z_e) (linear elasticity, per-element modulus).

CPU: the oracle's element kernel (orc_lelas_dkde_dot) inside a dense
adjoint is checked against central finite differences of dense solves on a
small cube, which pins its sign and scaling.
GPU: the autograd wrapper over the C ABI is checked against the oracle's
FP64 CPU adjoint and against fourth-order central differences (step 1e-2
in z) on 20 random elements, each to 1e-5 relative with no absolute floor,
at a test size and at the 1M-element configuration.  Forward, transposed
and difference solves are refined by residual-correction passes with
double-double residuals (svfsi_resid_dd_dev)."""
import sys

import numpy as np
import pytest

from svfsi_amd import problems

E0 = 1e6


def dense(Val, rp, cp, nNo, dof):
    K = np.zeros((nNo * dof, nNo * dof))
    for i in range(nNo):
        for k in range(rp[i], rp[i + 1]):
            j = cp[k]
            K[i * dof:(i + 1) * dof, j * dof:(j + 1) * dof] = Val[k].reshape(dof, dof)
    return K


def transpose_blocks(Val, rp, cp, nNo, dof):
    """Block-CSR values of K^T on the same (symmetric) pattern."""
    rows = np.repeat(np.arange(nNo), np.diff(rp))
    key = rows.astype(np.int64) * nNo + cp
    tkey = cp.astype(np.int64) * nNo + rows
    t = np.searchsorted(key, tkey)
    assert np.array_equal(key[t], tkey)
    return Val[t].reshape(-1, dof, dof).transpose(0, 2, 1).reshape(-1, dof * dof).copy()


def test_oracle_adjoint_vs_dense_fd(orc):
    prob = problems.lelas_adjoint_cube(3, sigma=0.3, ell=0.3)
    m, dof = prob.mesh, 3
    rp, cp = orc.lhsa(m.nNo, m.ien)
    free = prob.dir_mask().reshape(-1) == 1.0
    z = np.log(prob.prop_el / E0)
    rng = np.random.default_rng(7)
    u_obs = rng.normal(size=m.nNo * dof) * 1e-4
    u_obs[~free] = 0.0

    def solve(zz):
        prob.prop_el = E0 * np.exp(zz)
        V, R = orc.assemble(prob, rp, cp)
        K = dense(V, rp, cp, m.nNo, dof)
        u = np.zeros(m.nNo * dof)
        u[free] = np.linalg.solve(K[np.ix_(free, free)], R.reshape(-1)[free])
        return u, K

    def loss(zz):
        u, _ = solve(zz)
        return 0.5 * np.sum((u - u_obs) ** 2)

    u, K = solve(z)
    lam = np.zeros_like(u)
    lam[free] = np.linalg.solve(K[np.ix_(free, free)].T, (u - u_obs)[free])
    g = orc.lelas_dkde_dot(prob, lam.reshape(-1, 3), u.reshape(-1, 3)) * prob.prop_el
    h = 1e-5
    for e in rng.choice(m.nEl, 8, replace=False):
        zp, zm = z.copy(), z.copy()
        zp[e] += h
        zm[e] -= h
        fd = (loss(zp) - loss(zm)) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * np.abs(g).max(), (e, fd, g[e])


def _setup(n, relTol):
    import torch

    from svfsi_amd.adjoint import LelasAdjoint

    prob = problems.lelas_adjoint_cube(n, relTol=relTol)
    z = torch.tensor(np.log(prob.prop_el / E0), device="cuda")
    adj = LelasAdjoint(prob, passes=3)
    E_obs = E0 * np.exp(problems.random_field(prob.mesh.x[prob.mesh.ien].mean(axis=1), 2, 0.3, 0.1))
    u_obs, _ = adj.forward(torch.tensor(E_obs, device="cuda"), keep=False)
    return prob, adj, z, u_obs


def _fd_check(adj, z, u, u_obs, g, elems, h=1e-2, rtol=1e-5):
    """Fourth-order central differences of L in z_e,
    fd = [8 (D(h) - D(-h)) - (D(2h) - D(-2h))] / (12 h), with
    D(t) = L(z + t e_k) - L(z) = 0.5 d . (2 (u - u_obs) + d) and
    d = K(t)^-1 [(f(t) - f0) - (K(t) - K0) u] (LelasAdjoint.difference):
    the right-hand sides are differences of the assembled systems, nonzero
    only on the changed element's rows, and the base solve's own residual
    is left out exactly as the adjoint gradient, taken at the computed u,
    leaves it out; D is formed without cancellation.  A difference quotient
    divides the rounding of the computed L by the step, so the step is
    large (h = 1e-2) and the O(h^4) stencil keeps truncation below 1e-7.
    Returns the largest |fd - g_e| / (rtol |fd|), with no absolute floor."""
    import torch

    adj.prob.ls = dict(adj.prob.ls, relTol=1e-12)
    worst = 0.0
    for e in elems:
        D = {}
        for t in (h, -h, 2 * h, -2 * h):
            zz = z.clone()
            zz[e] += t
            d = adj.difference(E0 * torch.exp(zz), u)
            D[t] = 0.5 * torch.sum(d * (2 * (u - u_obs) + d)).item()
        fd = (8.0 * (D[h] - D[-h]) - (D[2 * h] - D[-2 * h])) / (12.0 * h)
        ge = g[e].item()
        ratio = abs(fd - ge) / (rtol * abs(fd))
        worst = max(worst, ratio)
        print("fd", int(e), fd, ge, "rel", abs(fd - ge) / max(abs(fd), 1e-300), "|g_e|/max",
              abs(ge) / g.abs().max().item(), flush=True)
    return worst


@pytest.mark.gpu
def test_adjoint_matches_cpu_adjoint_and_fd(gpu, orc):
    import torch

    from adjoint_fn import LelasSolve

    prob, adj, z, u_obs = _setup(8, 1e-12)
    m = prob.mesh
    zr = z.clone().requires_grad_(True)
    u = LelasSolve.apply(zr, adj, E0)
    L = 0.5 * torch.sum((u - u_obs) ** 2)
    L.backward()
    g = zr.grad.detach()
    # FP64 CPU adjoint (oracle): forward solve, transposed solve, element kernel
    rp, cp = orc.lhsa(m.nNo, m.ien)
    V, R = orc.assemble(prob, rp, cp)
    mask = prob.dir_mask()
    u_o, _, so = orc.fsils_solve(rp, cp, 3, mask, prob.ls, R, V)
    uo = u_obs.cpu().numpy()
    lam_o, _, sl = orc.fsils_solve(rp, cp, 3, mask, prob.ls, u_o - uo, transpose_blocks(V, rp, cp, m.nNo, 3))
    g_o = orc.lelas_dkde_dot(prob, lam_o, u_o) * prob.prop_el
    assert so.suc == 1 and sl.suc == 1
    gn = g.cpu().numpy()
    print("cpu adjoint relerr", np.abs(gn - g_o).max() / np.abs(g_o).max(), flush=True)
    assert np.abs(gn - g_o).max() <= 1e-7 * np.abs(g_o).max()
    assert np.abs(u.detach().cpu().numpy() - u_o).max() <= 1e-9 * np.abs(u_o).max()
    elems = np.random.default_rng(11).choice(m.nEl, 20, replace=False)
    assert _fd_check(adj, z, u.detach(), u_obs, g, elems) <= 1.0


@pytest.mark.gpu
def test_adjoint_fd_1m_elements(gpu):
    import torch

    from adjoint_fn import LelasSolve

    prob, adj, z, u_obs = _setup(55, 1e-12)
    print("nEl", prob.mesh.nEl, "nNo", prob.nNo, flush=True)
    zr = z.clone().requires_grad_(True)
    u = LelasSolve.apply(zr, adj, E0)
    L = 0.5 * torch.sum((u - u_obs) ** 2)
    L.backward()
    g = zr.grad.detach()
    elems = np.random.default_rng(12).choice(prob.mesh.nEl, 20, replace=False)
    assert _fd_check(adj, z, u.detach(), u_obs, g, elems) <= 1.0
