"""Taylor-Hood (P2P1, nFs = 2) FLUID assembly on the GPU (k_asm_fluid_th
and THOOD_ValRC) against the oracle, whose restatement the reference's own
output pins (tests/test_ref_golden.py, ref_fluid_th10_*): jittered tet10
cubes with random velocity, acceleration and pressure, the three viscosity
models and a per-element viscosity, FP64 to 1e-12 relative and bitwise
reproducible; the mid-edge pressure rows are the identity; and an FSILS
GMRES solve of the assembled system matches the oracle's; the mixed solve
(which picks the inner GMRES cycle without block Jacobi for a system with
zero pressure diagonals) reaches the FP64 criterion on the tet10 channel."""
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, mesh, problems
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu

VISC = {"const": (698, {}, 0.04),
        "cy": (697, dict(mu_o=0.56, lam=3.313, visc_a=2.0, visc_n=0.3568), 0.0345),
        "cass": (696, dict(mu_o=0.3, lam=0.5), 0.2)}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def th_problem(n, seed, visc="const", per_element=False):
    m = mesh.cube_tet10(n)
    rng = np.random.default_rng(seed)
    x = m.x.copy()
    h = 1.0 / n
    inner = np.all((x > 1e-12) & (x < 1 - 1e-12), axis=1)
    corners = np.unique(m.ien[:, :4])
    mv = np.zeros_like(x)
    mv[corners] = rng.uniform(-0.2 * h, 0.2 * h, (len(corners), 3))
    mv[~inner] = 0.0
    x[corners] += mv[corners]
    for k, (a, b) in enumerate(mesh._MID):
        x[m.ien[:, 4 + k]] = 0.5 * (x[m.ien[:, a]] + x[m.ien[:, b]])
    m = mesh.TetMesh(x=x, ien=m.ien, node_tags={})
    vt, vp, mu = VISC[visc]
    eq = dict(problems.gen_alpha(0.2, False), dt=0.01, rho=1.06, mu=mu, bstab=0.2, fb=(0.3, -0.2, 0.1),
              viscType=vt, **vp)
    Y = np.zeros((m.nNo, 4))
    Y[:, :3] = rng.uniform(-5.0, 5.0, (m.nNo, 3))
    Y[corners, 3] = rng.uniform(-1.0, 1.0, len(corners))
    A = rng.normal(size=(m.nNo, 4)) * 5.0
    x0 = np.where(m.x[:, 0] < 1e-12)[0].astype(np.int32)
    p = problems.Problem(name=f"th10_{visc}", phys=problems.SVFSI_PHYS_FLUID, dof=4, mesh=m, eq=eq, Ag=A, Yg=Y,
                         Dg=np.zeros_like(Y), dir_faces=[(x0, 3)], nFs=2,
                         ls=dict(LS_type=fsils.LS_TYPE_GMRES, relTol=1e-10, absTol=1e-30, mItr=100, sD=50))
    if per_element:
        p.prop_el = mu * np.exp(0.2 * rng.normal(size=m.nEl))
    return p


def rel_val(V, Vo):
    return np.abs(V - Vo).max() / np.abs(Vo[:, :15]).max()


@pytest.mark.parametrize("visc,gather", [("const", "1"), ("cy", "1"), ("cass", "1"), ("const", "0")])
def test_th_assembly_matches_oracle(gpu, orc, monkeypatch, visc, gather):
    """gather "1": one launch storing element blocks, summed per Val block in
    element order (the default); "0": coloured read-modify-write."""
    import torch

    monkeypatch.setenv("SVFSI_ASM_GATHER", gather)
    p = th_problem(4, 11, visc, per_element=(visc == "const"))
    say("Taylor-Hood tet10 assembly", visc, "nEl", p.mesh.nEl, "nNo", p.nNo)
    st = NewtonStep(p)
    st.assemble()
    torch.cuda.synchronize()
    V, R = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    Vo, Ro = orc.assemble(p, st.rowPtr - 1, st.colPtr - 1)
    ev, er = rel_val(V, Vo), np.abs(R - Ro).max() / np.abs(Ro).max()
    say(visc, "Val", ev, "R", er)
    assert ev < 1e-12 and er < 1e-12
    # bitwise reproducible (colour-ordered read-modify-write, no atomics)
    st.assemble()
    torch.cuda.synchronize()
    assert np.array_equal(st.Val.cpu().numpy(), V) and np.array_equal(st.R.cpu().numpy(), R)
    # THOOD_ValRC: mid-edge pressure rows are the identity, residual 0
    mids = np.unique(p.mesh.ien[:, 4:])
    rp, cp = st.rowPtr - 1, st.colPtr - 1
    for n in mids[:50]:
        assert R[n, 3] == 0.0
        for k in range(rp[n], rp[n + 1]):
            assert V[k, 15] == (1.0 if cp[k] == n else 0.0)


def test_th_gmres_solve_matches_oracle(gpu, orc):
    import torch

    p = th_problem(3, 12)
    st = NewtonStep(p)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve()
    torch.cuda.synchronize()
    x = st.R.cpu().numpy()
    xo, _, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, 4, p.dir_mask(), p.ls, R0, V0)
    err = np.abs(x - xo).max() / np.abs(xo).max()
    say("Taylor-Hood GMRES: itr", ls.RI.itr, "oracle", so.itr, "x rel err", err)
    assert ls.RI.suc == 1 and so.suc == 1
    assert abs(ls.RI.itr - so.itr) <= max(2, so.itr // 50)
    assert err < 1e-8


def test_th_needs_the_flag_on_tet10(gpu):
    from svfsi_amd._lib import SvfsiError

    p = th_problem(2, 13)
    st = NewtonStep(p)
    with pytest.raises(SvfsiError, match="P2P1"):
        st.dmesh.assemble(p.phys, p.eq, st.Ag, st.Yg, st.Dg, st.Val, st.R, quad=0)


def test_th_mixed_solve_reaches_fp64_tolerance(gpu, orc):
    import torch

    p = problems.fluid_th_cube(4)
    st = NewtonStep(p)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve(mixed=True)
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy()
    rp, cp = st.rowPtr - 1, st.colPtr - 1
    mask = p.dir_mask()
    _, Rs, W = orc.precond_diag(rp, cp, 4, V0, R0, mask)
    xo, Vs, so = orc.fsils_solve(rp, cp, 4, mask, p.ls, R0, V0)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(rp, cp, 4, Vs, y)
    say("Taylor-Hood mixed: outer", ls.RI.itr, "SpMVs", ls.RI.cM, "scaled residual",
        np.linalg.norm(r) / np.linalg.norm(Rs), "x rel err", np.abs(xg - xo).max() / np.abs(xo).max())
    assert ls.RI.suc == 1 and so.suc == 1
    assert np.all(np.isfinite(xg))
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
    assert np.abs(xg - xo).max() <= 1e-4 * np.abs(xo).max()
