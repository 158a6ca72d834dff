"""Viscosity adjoint of one linearised VMS fluid step: K(mu) du = R(mu),
mu_e = mu0 exp(z_e), loss L = c . du, the outlet flux of du weighted by
1 - r^2.

CPU: the oracle's complex-step element sensitivity (orc_fluid_dmu_dot) is
checked against central differences of the oracle's assembled K and R, and
the dense adjoint built on it against central differences of the dense
solve.
GPU: the dual-number kernel (svfsi_fluid_dmu_dot_dev) against the oracle;
the autograd wrapper over the C ABI against the oracle's FP64 CPU adjoint
and fourth-order central differences (step 4e-2 in z = log(mu / mu0)) on
20 elements drawn at random among those with |dL/dz_e| >= 1e-3 max|dL/dz|
(the rest change the outlet-flux loss by less than the rounding of L), each
to 1e-5 relative with no absolute floor.  Every solve (forward, transposed,
FD differences) is refined by three residual-correction passes with
double-double residuals: K's condition number (~4e8) leaves one solve to
1e-12 only ~4e-4 accurate."""
import numpy as np
import pytest

from svfsi_amd import problems

MU0 = 0.04


def outlet_flux_weights(prob):
    """c with c . du = integral over the outlet of (1 - r^2) du . n (tri3,
    integral of N_a: area / 3; outlet normal +z).  The weight measures the
    profile's shape: the plain outlet flux of an increment is pinned by
    continuity (the inlet and wall increments are zero) and is nearly
    independent of the viscosity."""
    m = prob.mesh
    c = np.zeros((m.nNo, 4))
    wr = 1.0 - (m.x[:, 0] ** 2 + m.x[:, 1] ** 2)
    for fien, _ in prob.bflow_faces:
        p = m.x[fien]
        area = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        for k in range(3):
            np.add.at(c[:, 2], fien[:, k], area / 3.0 * wr[fien[:, k]])
    c[prob.dir_mask() == 0.0] = 0.0
    return c


def fluid_problem(h, seed=5):
    prob = problems.fluid_cylinder(h=h)
    rng = np.random.default_rng(seed)
    prob.Ag[:, :3] = rng.normal(size=(prob.nNo, 3))
    prob.prop_el = MU0 * np.exp(0.1 * rng.normal(size=prob.mesh.nEl))
    return prob


def dense(Val, rp, cp, nNo, dof):
    K = np.zeros((nNo * dof, nNo * dof))
    for i in range(nNo):
        for k in range(rp[i], rp[i + 1]):
            K[i * dof:(i + 1) * dof, cp[k] * dof:(cp[k] + 1) * dof] = Val[k].reshape(dof, dof)
    return K


def test_oracle_dmu_matches_fd_of_assembly(orc):
    prob = fluid_problem(0.6)
    m = prob.mesh
    rp, cp = orc.lhsa(m.nNo, m.ien)
    rng = np.random.default_rng(1)
    lam = rng.normal(size=(m.nNo, 4))
    du = rng.normal(size=(m.nNo, 4))
    g = orc.fluid_dmu_dot(prob, lam, du)
    mu = prob.prop_el.copy()
    for e in rng.choice(m.nEl, 6, replace=False):
        h = 1e-4 * mu[e]  # O(h^2) truncation ~1e-8; rounding of |R| ~ 1e3 stays below 1e-7
        out = []
        for s in (1, -1):
            prob.prop_el = mu.copy()
            prob.prop_el[e] += s * h
            V, R = orc.assemble(prob, rp, cp)
            Kd = dense(V, rp, cp, m.nNo, 4)
            out.append(R.ravel() - Kd @ du.ravel())
        fd = lam.ravel() @ (out[0] - out[1]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * abs(g[e]) + 1e-9 * np.abs(g).max(), (e, fd, g[e])
    prob.prop_el = mu


def test_oracle_dense_adjoint_vs_fd(orc):
    prob = fluid_problem(0.5)
    m = prob.mesh
    rp, cp = orc.lhsa(m.nNo, m.ien)
    free = prob.dir_mask().ravel() == 1.0
    c = outlet_flux_weights(prob).ravel()
    mu = prob.prop_el.copy()

    def solve(mm):
        prob.prop_el = mm
        V, R = orc.assemble(prob, rp, cp)
        K = dense(V, rp, cp, m.nNo, 4)
        x = np.zeros(m.nNo * 4)
        x[free] = np.linalg.solve(K[np.ix_(free, free)], R.ravel()[free])
        return x, K

    x, K = solve(mu)
    lam = np.zeros_like(x)
    lam[free] = np.linalg.solve(K[np.ix_(free, free)].T, c[free])
    prob.prop_el = mu
    g = orc.fluid_dmu_dot(prob, lam.reshape(-1, 4), x.reshape(-1, 4))
    # the dense FD of L (|L| ~ 4e4, K's condition ~ 4e8) resolves only the
    # sensitive elements: the 5 largest |g|
    for e in np.argsort(-np.abs(g))[:5]:
        h = 1e-3 * mu[e]
        L = []
        for s in (1, -1):
            mm = mu.copy()
            mm[e] += s * h
            L.append(c @ solve(mm)[0])
        fd = (L[0] - L[1]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-4 * abs(g[e]), (e, fd, g[e])


@pytest.mark.gpu
def test_fluid_dmu_kernel_parity(gpu, orc):
    import torch

    from svfsi_amd.step import NewtonStep

    prob = fluid_problem(0.4)
    st = NewtonStep(prob)
    rng = np.random.default_rng(3)
    lam = rng.normal(size=(prob.nNo, 4))
    du = rng.normal(size=(prob.nNo, 4))
    dl, dd = (torch.from_numpy(a).to(gpu) for a in (lam, du))
    g = torch.zeros(prob.mesh.nEl, dtype=torch.float64, device=gpu)
    st.dmesh.fluid_dmu_dot(prob.eq, st.Ag, st.Yg, st.prop, dl, dd, g)
    torch.cuda.synchronize()
    go = orc.fluid_dmu_dot(prob, lam, du)
    gg = g.cpu().numpy()
    assert np.abs(gg - go).max() <= 1e-12 * np.abs(go).max()


def _fd_check(adj, z, du, c, g, elems, h=4e-2, rtol=1e-5):
    """Fourth-order central differences of L = c . du in z_e,
    fd = [8 (D(h) - D(-h)) - (D(2h) - D(-2h))] / (12 h), with
    D(t) = L(z + t e_k) - L(z) = c . d(t) and d(t) = K(t)^-1 [(R(t) - R0) -
    (K(t) - K0) du] (FluidViscosityAdjoint.difference: no cancellation in
    forming D).  Returns the largest |fd - g_e| / (rtol |fd|).

    The step: the computed L(z) carries the rounding of the FP64 assembly
    (the pressure increments reach 3e8, so eps |K_e| |du_e| is not small
    against h dK_e/dz du_e), and a difference quotient divides it by h.
    Second-order differences at h = 1e-4 miss g_e by up to 6e-4 on elements
    with |g_e| ~ 1e-3 max|g|, and the miss grows as 1/h (8e-4 at 5e-5, 7e-5
    at 4e-4) -- the same digits with dense LU solves on the CPU, so it is
    the function's own rounding, not the solves'.  The fourth-order stencil
    at h = 4e-2 has O(h^4) truncation and divides the rounding by a step
    400 times larger: 1.6e-6 worst on the CPU over 15 such elements."""
    import torch

    adj.prob.ls = dict(adj.prob.ls, relTol=1e-12, mItr=max(adj.prob.ls.get("mItr", 0), 200))
    worst = 0.0
    for e in elems:
        D = {}
        for t in (h, -h, 2 * h, -2 * h):
            zz = z.clone()
            zz[e] += t
            D[t] = torch.sum(c * adj.difference(MU0 * torch.exp(zz), du)).item()
        fd = (8.0 * (D[h] - D[-h]) - (D[2 * h] - D[-2 * h])) / (12.0 * h)
        ge = g[e].item()
        ratio = abs(fd - ge) / (rtol * abs(fd))
        worst = max(worst, ratio)
        print("fd", int(e), fd, ge, "rel", abs(fd - ge) / max(abs(fd), 1e-300), "|g_e|/max",
              abs(ge) / g.abs().max().item(), flush=True)
    return worst


def resolvable_sample(g, n, seed, floor=1e-3):
    """n elements drawn at random (seeded) among those whose |dL/dz_e| is at
    least floor * max|dL/dz|.  The loss is a weighted outlet flux: most
    elements of the long cylinder change it by less than the rounding of L
    itself (|g_e| down to 1e-10 of the largest), so a uniform draw picks
    elements whose difference quotient is rounding."""
    gg = g.abs().cpu().numpy()
    cand = np.nonzero(gg >= floor * gg.max())[0]
    return np.random.default_rng(seed).choice(cand, min(n, len(cand)), replace=False)


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_viscosity_adjoint_vs_cpu_adjoint_and_fd(gpu, orc, mixed):
    import torch

    from adjoint_fn import FluidViscositySolve
    from svfsi_amd.adjoint import FluidViscosityAdjoint

    prob = fluid_problem(0.5)
    prob.ls = dict(prob.ls, relTol=1e-12, mItr=5000)
    m = prob.mesh
    adj = FluidViscosityAdjoint(prob, mixed=mixed, passes=3)
    z = torch.tensor(np.log(prob.prop_el / MU0), device=gpu)
    c = torch.tensor(outlet_flux_weights(prob), device=gpu)
    zr = z.clone().requires_grad_(True)
    du = FluidViscositySolve.apply(zr, adj, MU0)
    L = torch.sum(c * du)
    L.backward()
    g = zr.grad.detach()
    # FP64 CPU adjoint: oracle forward and transposed GMRES, complex-step element terms
    rp, cp = orc.lhsa(m.nNo, m.ien)
    V, R = orc.assemble(prob, rp, cp)
    mask = prob.dir_mask()
    x_o, _, so = orc.fsils_solve(rp, cp, 4, mask, prob.ls, R, V)
    from test_adjoint import transpose_blocks

    lam_o, _, sl = orc.fsils_solve(rp, cp, 4, mask, prob.ls, c.cpu().numpy(), transpose_blocks(V, rp, cp, m.nNo, 4))
    assert so.suc == 1 and sl.suc == 1
    g_o = orc.fluid_dmu_dot(prob, lam_o, x_o) * prob.prop_el
    gn = g.cpu().numpy()
    err = np.abs(gn - g_o).max() / np.abs(g_o).max()
    print("cpu adjoint relerr", err, flush=True)
    assert err <= 1e-7
    elems = resolvable_sample(g, 20, 13)
    assert _fd_check(adj, z, du.detach(), c, g, elems) <= 1.0


@pytest.mark.gpu
def test_viscosity_adjoint_large_cylinder(gpu):
    """The config-5 adjoint at full size: the ~3.2M-tet cylinder (h = 0.042),
    mu_e = 0.04 exp(z_e) with z ~ N(0, 0.1^2) i.i.d., loss = the weighted
    outlet flux of the linearised step, mixed-precision forward and transposed
    solves.  Checked by central differences (relative step 1e-4 in mu) on the
    five elements of largest |dL/dz| and five drawn at random among those
    with |dL/dz_e| >= 1e-3 max|dL/dz|, each to 1e-5 relative."""
    import time

    import torch

    from adjoint_fn import FluidViscositySolve
    from svfsi_amd.adjoint import FluidViscosityAdjoint

    prob = fluid_problem(0.042)
    prob.ls = dict(prob.ls, relTol=1e-12, mItr=200, sD=100)
    m = prob.mesh
    print("nEl", m.nEl, "nNo", m.nNo, flush=True)
    adj = FluidViscosityAdjoint(prob, mixed=True, passes=3)
    z = torch.tensor(np.log(prob.prop_el / MU0), device=gpu)
    c = torch.tensor(outlet_flux_weights(prob), device=gpu)
    zr = z.clone().requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    du = FluidViscositySolve.apply(zr, adj, MU0)
    L = torch.sum(c * du)
    L.backward()
    torch.cuda.synchronize()
    print(f"forward + adjoint {time.perf_counter() - t0:.2f} s", flush=True)
    g = zr.grad.detach()
    assert torch.isfinite(g).all()
    top = torch.argsort(g.abs(), descending=True)[:5].cpu().numpy()
    rnd = resolvable_sample(g, 5, 17)
    assert _fd_check(adj, z, du.detach(), c, g, np.concatenate([top, rnd])) <= 1.0
