"""Pins of the oracle's non-Newtonian viscosity (GETVISCOSITY,
svFSI/FLUID.f:1690-1739, and its use in FLUID3D_M/C, FLUID.f:329-373,
467-545, 1026-1065):

- Carreau-Yasuda against the published model mu = mu_inf + (mu_0 - mu_inf)
  (1 + (lam g)^a)^((n-1)/a) (limits g -> 0 and g -> inf, the power-law
  slope), n = 1 reducing to the constant mu_0;
- Casson against its law sqrt(tau) = sqrt(tau_y) + sqrt(eta g) with
  mu_o = sqrt(tau_y), mu_i = sqrt(eta), constant below the threshold;
- d mu / d gamma against central differences (Carreau-Yasuda; Casson's
  value in the reference is -2 times the derivative of its viscosity, kept
  as the reference computes it);
- tet4 elements have a constant velocity gradient, so a non-Newtonian
  assembly equals the Newtonian one with the per-element viscosity
  mu(gamma_e) in the residual, the continuity rows and the pressure
  columns; in the velocity blocks it adds exactly
  af gam dt vol_e (mu'(gamma_e)/gamma_e) g_a,i g_b,j with g_a = es Nx_a,
  the derivative of the Galerkin stress mu(gamma) es through gamma."""
import numpy as np
import pytest

from svfsi_amd import problems

CY = dict(viscType=697, mu=0.0345, mu_o=0.56, lam=3.313, visc_a=2.0, visc_n=0.3568)
CASS = dict(viscType=696, mu=0.187, mu_o=0.2, lam=0.5)


def test_carreau_yasuda_closed_forms(orc):
    mu_inf, mu_0, lam, a, n = CY["mu"], CY["mu_o"], CY["lam"], CY["visc_a"], CY["visc_n"]
    mu, _, _, _ = orc.getviscosity(CY, 0.0)
    assert mu == pytest.approx(mu_0, rel=1e-15)
    mu, _, _, _ = orc.getviscosity(CY, 1e12)
    assert mu == pytest.approx(mu_inf, rel=1e-6)
    for g in (1e-3, 0.1, 1.0, 37.0, 1e3):
        mu, mus, _, gout = orc.getviscosity(CY, g)
        assert mu == pytest.approx(mu_inf + (mu_0 - mu_inf) * (1 + (lam * g) ** a) ** ((n - 1) / a), rel=1e-13)
        assert mus == mu_inf and gout == g
    # power-law regime: d log(mu - mu_inf) / d log g -> n - 1
    g1, g2 = 1e6, 2e6
    m1 = orc.getviscosity(CY, g1)[0] - mu_inf
    m2 = orc.getviscosity(CY, g2)[0] - mu_inf
    assert np.log(m2 / m1) / np.log(g2 / g1) == pytest.approx(n - 1, rel=1e-6)
    # n = 1: constant mu_0
    for g in (0.0, 3.0, 1e4):
        assert orc.getviscosity(dict(CY, visc_n=1.0), g)[0] == pytest.approx(mu_0, rel=1e-15)


def test_casson_closed_forms(orc):
    eta, tau_y = CASS["mu"] ** 2, CASS["mu_o"] ** 2
    for g in (0.5, 0.7, 3.0, 250.0):
        mu, mus, _, gout = orc.getviscosity(CASS, g)
        assert np.sqrt(mu * g) == pytest.approx(np.sqrt(tau_y) + np.sqrt(eta * g), rel=1e-14)
        assert mus == pytest.approx(eta, rel=1e-15) and gout == g
    mlam = orc.getviscosity(CASS, CASS["lam"])[0]
    for g in (0.0, 0.1, 0.49):
        mu, _, _, gout = orc.getviscosity(CASS, g)
        assert mu == mlam and gout == CASS["lam"]


@pytest.mark.parametrize("model", ["cy", "cass"])
def test_viscosity_derivative_matches_central_differences(orc, model):
    """Carreau-Yasuda: mu_x is d mu / d gamma.  Casson: the reference's
    mu_x = 2 mu_o' (mu_o' + mu_i) / gamma (FLUID.f:1733, mu_o' = mu_o /
    sqrt(gamma)) is -2 times the derivative of its own
    mu = (mu_i + mu_o')^2, which is -mu_o' (mu_o' + mu_i) / gamma; the
    restatement keeps the reference's value, and below the threshold it is
    the value at the raised rate lam."""
    eq = CY if model == "cy" else CASS
    factor = 1.0 if model == "cy" else -2.0
    for g in (0.6, 2.0, 45.0):
        _, _, mux, _ = orc.getviscosity(eq, g)
        h = 1e-6 * g
        fd = (orc.getviscosity(eq, g + h)[0] - orc.getviscosity(eq, g - h)[0]) / (2 * h)
        assert mux == pytest.approx(factor * fd, rel=1e-7)
    if model == "cass":
        assert orc.getviscosity(eq, 0.1)[2] == orc.getviscosity(eq, eq["lam"])[2]


def element_gamma(m, Y):
    """Per-element shear rate of the linear velocity field: grad u from the
    affine map of each tet, es = grad u + grad u^T, gamma = sqrt(es:es / 2);
    also the volume and the shape-function gradients."""
    x = m.x[m.ien]                       # (nEl, 4, 3)
    J = (x[:, 1:] - x[:, :1])            # rows: edges from node 0
    Jinv = np.linalg.inv(J)              # grad of the barycentric coords 1..3
    dN = np.concatenate([-Jinv.sum(axis=2, keepdims=True), Jinv], axis=2)  # (nEl, 3, 4): dN_a/dx_i
    u = Y[m.ien][:, :, :3]               # (nEl, 4, 3)
    gu = np.einsum("eia,eac->eic", dN, u)  # du_c/dx_i
    es = gu + np.transpose(gu, (0, 2, 1))
    gam = np.sqrt(0.5 * np.einsum("eij,eij->e", es, es))
    vol = np.abs(np.linalg.det(J)) / 6.0
    return gam, es, dN, vol


@pytest.mark.parametrize("model", ["cy", "cass"])
def test_nonnewtonian_assembly_against_per_element_newtonian(orc, model):
    visc = CY if model == "cy" else CASS
    prob = problems.fluid_cylinder(h=0.7)
    rng = np.random.default_rng(5)
    # velocities whose shear rates span the model's regimes (Casson: both
    # sides of its threshold lam = 0.5)
    prob.Yg[:, :3] = (2.0 if model == "cy" else 0.1) * rng.normal(size=(prob.nNo, 3))
    m = prob.mesh
    rp, cp = orc.lhsa(m.nNo, m.ien)
    prob.eq = dict(prob.eq, **visc)
    Vn, Rn = orc.assemble(prob, rp, cp)
    gam, es, dN, vol = element_gamma(m, prob.Yg)
    mu_e = np.array([orc.getviscosity(visc, g)[0] for g in gam])
    mux = np.array([orc.getviscosity(visc, g)[2] for g in gam])
    geff = np.array([orc.getviscosity(visc, g)[3] for g in gam])  # Casson raises gamma to lam
    if model == "cass":
        assert (gam < visc["lam"]).any() and (gam > visc["lam"]).any()
    newt = problems.fluid_cylinder(h=0.7)
    newt.Yg = prob.Yg
    newt.eq = dict(prob.eq, viscType=698)
    newt.prop_el = mu_e
    V0, R0 = orc.assemble(newt, rp, cp)
    assert np.abs(Rn - R0).max() <= 1e-13 * np.abs(R0).max()
    dV = (Vn - V0).reshape(-1, 4, 4)
    # continuity rows and pressure columns unchanged
    assert np.abs(dV[:, 3, :]).max() <= 1e-13 * np.abs(V0).max()
    assert np.abs(dV[:, :3, 3]).max() <= 1e-13 * np.abs(V0).max()
    # velocity blocks: the rank-one term per element
    eq = prob.eq
    c = eq["af"] * eq["gam"] * eq["dt"]
    mug = np.where(geff > 0, mux / np.where(geff > 0, geff, 1.0), 0.0)
    g = np.einsum("ejc,eja->eac", es, dN)  # g[e, a, c] = sum_j es(j,c) dN_a/dx_j
    expect = np.zeros_like(dV[:, :3, :3])
    col = {}
    for i in range(m.nNo):
        for j in range(rp[i], rp[i + 1]):
            col[(i, cp[j])] = j
    for e in range(m.nEl):
        coef = c * vol[e] * mug[e]
        for a in range(4):
            for b in range(4):
                expect[col[(m.ien[e, a], m.ien[e, b])]] += coef * np.outer(g[e, a], g[e, b])
    scale = np.abs(V0[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]]).max()
    assert np.abs(expect).max() > 1e-4 * scale  # the term is not negligible here
    assert np.abs(dV[:, :3, :3] - expect).max() <= 1e-11 * scale
