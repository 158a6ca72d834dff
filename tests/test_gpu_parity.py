"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer/index work is compared exactly, FP64 kernels to
1e-12 relative (global, max-norm), FP32 kernels to 1e-5 relative."""
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, mesh, problems
from svfsi_amd._lib import SVFSI_STRUCT_MFMA32, SVFSI_STRUCT_MFMA64, SVFSI_STRUCT_VALU
from svfsi_amd.assembly import DeviceMesh
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def T(a, gpu, dtype=None):
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(gpu)


def relerr(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def cube(orc):
    m = mesh.cube_tet4(5)
    rp, cp = fsils.lhsa(m.nNo, m.ien)
    return m, rp, cp


@pytest.mark.parametrize("dof", [3, 4])
def test_spmv_fp64_fp32_transpose(gpu, orc, cube, dof):
    import torch

    m, rp, cp = cube
    say("spmv", dof, m.nNo, len(cp))
    lhs = fsils.fsils_lhs_create(m.nNo, len(cp), rp, cp)
    rng = np.random.default_rng(dof)
    Val = rng.normal(size=(len(cp), dof * dof))
    U = rng.normal(size=(m.nNo, dof))
    ref = orc.spmv(rp - 1, cp - 1, dof, Val, U)
    out = torch.zeros((m.nNo, dof), dtype=torch.float64, device=gpu)
    dV, dU = T(Val, gpu), T(U, gpu)  # keep the tensors alive across the calls
    fsils.call("svfsi_spmv_dev", lhs.handle, dof, fsils.dev_ptr(dV), fsils.dev_ptr(dU),
               fsils.dev_ptr(out), fsils.cur_stream())
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-13
    # transpose product
    reft = orc.spmv_t(rp - 1, cp - 1, dof, Val, U)
    fsils.call("svfsi_spmv_t_dev", lhs.handle, dof, fsils.dev_ptr(dV), fsils.dev_ptr(dU),
               fsils.dev_ptr(out), fsils.cur_stream())
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), reft) < 1e-13
    # FP32 values/vectors: compare with the FP64 product of the rounded inputs
    V32 = Val.astype(np.float32)
    U32 = U.astype(np.float32)
    ref32 = orc.spmv(rp - 1, cp - 1, dof, V32.astype(float), U32.astype(float))
    o32 = torch.zeros((m.nNo, dof), dtype=torch.float32, device=gpu)
    dV32, dU32 = T(V32, gpu), T(U32, gpu)
    fsils.call("svfsi_spmv_f32_dev", lhs.handle, dof, fsils.dev_ptr(dV32), fsils.dev_ptr(dU32),
               fsils.dev_ptr(o32), fsils.cur_stream())
    torch.cuda.synchronize()
    assert relerr(o32.cpu().numpy(), ref32) < 1e-5
    lhs.free()


def test_dot(gpu):
    import torch

    from svfsi_amd._lib import lib
    import ctypes as C

    rng = np.random.default_rng(0)
    for n in (1, 255, 1000, 123457):
        a, b = rng.normal(size=n), rng.normal(size=n)
        out = C.c_double()
        da, db = T(a, gpu), T(b, gpu)
        fsils.call("svfsi_dot_dev", n, fsils.dev_ptr(da), fsils.dev_ptr(db), C.byref(out),
                   fsils.cur_stream())
        assert abs(out.value - a @ b) < 1e-12 * np.abs(a * b).sum()


def test_precond(gpu, orc, cube):
    import torch

    m, rp, cp = cube
    dof = 3
    lhs = fsils.fsils_lhs_create(m.nNo, len(cp), rp, cp, nFaces=1)
    fixed = np.where(m.x[:, 0] == 0.0)[0]
    fsils.fsils_bc_create(lhs, 1, len(fixed), 3, fsils.BC_TYPE_Dir, fixed + 1)
    rng = np.random.default_rng(1)
    Val = rng.normal(size=(len(cp), 9))
    R = rng.normal(size=(m.nNo, 3))
    mask = np.ones((m.nNo, 3))
    mask[fixed] = 0
    V2, R2, W2 = orc.precond_diag(rp - 1, cp - 1, dof, Val, R, mask)
    dV, dR = T(Val, gpu), T(R, gpu)
    dW = torch.zeros((m.nNo, 3), dtype=torch.float64, device=gpu)
    fsils.call("svfsi_precond_diag_dev", lhs.handle, dof, fsils.dev_ptr(dV), fsils.dev_ptr(dR), fsils.dev_ptr(dW),
               None, fsils.cur_stream())
    assert np.array_equal(dW.cpu().numpy(), W2)
    assert np.array_equal(dR.cpu().numpy(), R2)
    assert relerr(dV.cpu().numpy(), V2) < 1e-15
    lhs.free()


def _assembled(prob, orc):
    st = NewtonStep(prob)
    st.assemble()
    import torch

    torch.cuda.synchronize()
    Vg, Rg = st.Val.cpu().numpy(), st.R.cpu().numpy()
    Vo, Ro = orc.assemble(prob, st.rowPtr - 1, st.colPtr - 1)
    return st, Vg, Rg, Vo, Ro


@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_precond_rcs_parity(gpu, orc, which):
    """PRECONDRCS on the device against the oracle: scaled matrix and
    right-hand side and both factor vectors to 1e-14 (max and sqrt are exact
    or correctly rounded, the products are taken in the same order)."""
    import torch

    from svfsi_amd._lib import call

    prob = problems.fluid_cylinder(h=0.5) if which == "fluid" else problems.lelas_cube(5)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    W1 = torch.zeros_like(st.R)
    W2 = torch.zeros_like(st.R)
    call("svfsi_precond_rcs_dev", st.lhs.handle, prob.dof, fsils.dev_ptr(st.Val), fsils.dev_ptr(st.R),
         fsils.dev_ptr(W1), fsils.dev_ptr(W2), None, fsils.cur_stream())
    Vo, Ro, W1o, W2o = orc.precond_rcs(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, prob.dir_mask())
    say("rcs parity", which, relerr(st.Val.cpu().numpy(), Vo), relerr(W2.cpu().numpy(), W2o))
    assert relerr(st.Val.cpu().numpy(), Vo) < 1e-14
    assert relerr(st.R.cpu().numpy(), Ro) < 1e-14
    assert relerr(W1.cpu().numpy(), W1o) < 1e-14
    assert relerr(W2.cpu().numpy(), W2o) < 1e-14


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_solve_with_rcs(gpu, orc, which, mixed):
    """FSILS_SOLVE with PRECOND_RCS (GMRES for the fluid, CG for LELAS), FP64
    and mixed paths, against the oracle's solve with the same
    preconditioner."""
    prob = problems.fluid_cylinder(h=0.5) if which == "fluid" else problems.lelas_cube(5)
    prob.ls = dict(prob.ls, prec=709)
    st = NewtonStep(prob)
    st.assemble()
    import torch

    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy()
    xo, _, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), prob.ls, R0, V0, prec=709)
    say("rcs solve", which, "mixed", mixed, "itr", ls.RI.itr, "cpu", so.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and so.suc == 1
    if not mixed:
        assert abs(ls.RI.itr - so.itr) <= max(3, so.itr // 50), (ls.RI.itr, so.itr)
    assert relerr(xg, xo) < 1e-5


def test_lelas_assembly_parity_and_determinism(gpu, orc):
    prob = problems.lelas_cube(6)
    rng = np.random.default_rng(3)
    prob.Dg = rng.normal(size=prob.Dg.shape) * 1e-3
    prob.Ag = rng.normal(size=prob.Ag.shape)
    say("lelas assembly", prob.nNo)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12
    st.assemble()
    import torch

    torch.cuda.synchronize()
    assert np.array_equal(st.Val.cpu().numpy(), Vg)  # bitwise reproducible
    assert np.array_equal(st.R.cpu().numpy(), Rg)


def test_lelas_per_element_modulus(gpu, orc):
    prob = problems.lelas_cube(4)
    prob.prop_el = 1e6 * np.exp(np.random.default_rng(2).normal(size=prob.mesh.nEl) * 0.3)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    assert relerr(Vg, Vo) < 1e-12


def test_fluid_assembly_parity(gpu, orc):
    prob = problems.fluid_cylinder(h=0.4)
    rng = np.random.default_rng(4)
    prob.Ag[:, :3] = rng.normal(size=(prob.nNo, 3))
    prob.Yg[:, 3] = rng.normal(size=prob.nNo)
    say("fluid assembly", prob.nNo, prob.mesh.nEl)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12


VISC = {
    "cy": dict(viscType=697, mu=0.0345, mu_o=0.56, lam=3.313, visc_a=2.0, visc_n=0.3568),
    "cass": dict(viscType=696, mu=0.187, mu_o=0.2, lam=0.5),
}


@pytest.mark.parametrize("model", ["cy", "cass"])
def test_fluid_assembly_nonnewtonian(gpu, orc, model):
    """Carreau-Yasuda and Casson viscosity (GETVISCOSITY) in the fluid
    element kernel against the oracle to 1e-12, with shear rates on both
    sides of the models' transitions."""
    prob = problems.fluid_cylinder(h=0.4)
    rng = np.random.default_rng(6)
    prob.Yg[:, :3] = (2.0 if model == "cy" else 0.1) * rng.normal(size=(prob.nNo, 3))
    prob.Ag[:, :3] = rng.normal(size=(prob.nNo, 3))
    prob.eq = dict(prob.eq, **VISC[model])
    say("fluid assembly", model, prob.nNo, prob.mesh.nEl)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    newt = problems.fluid_cylinder(h=0.4)
    newt.Yg, newt.Ag = prob.Yg, prob.Ag
    V0, _ = orc.assemble(newt, st.rowPtr - 1, st.colPtr - 1)
    say("relerr", relerr(Vg, Vo), relerr(Rg, Ro), "vs Newtonian", relerr(Vo, V0))
    assert relerr(Vo, V0) > 1e-3  # the model changed the matrix
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12


def test_viscosity_model_errors(gpu):
    """An unknown viscType is refused, and so is the viscosity sensitivity
    (defined for the constant model) with a non-Newtonian model."""
    import torch

    from svfsi_amd._lib import SvfsiError

    prob = problems.fluid_cylinder(h=0.8)
    prob.eq = dict(prob.eq, viscType=695)
    st = NewtonStep(prob)
    with pytest.raises(SvfsiError, match="viscType"):
        st.assemble()
    eq = dict(prob.eq, **VISC["cy"])
    n = prob.nNo
    z = torch.zeros((n, 4), dtype=torch.float64, device=gpu)
    g = torch.zeros(prob.mesh.nEl, dtype=torch.float64, device=gpu)
    with pytest.raises(SvfsiError, match="constant"):
        st.dmesh.fluid_dmu_dot(eq, st.Ag, st.Yg, None, z, z, g)


@pytest.mark.parametrize("kern", [SVFSI_STRUCT_VALU, SVFSI_STRUCT_MFMA64])
@pytest.mark.parametrize("quad", [0, 1])
def test_struct_assembly_parity(gpu, orc, quad, kern):
    """STRUCT tet10 FP64 assembly (vector-unit and FP64-MFMA kernels) within
    1e-12 of the oracle and bitwise reproducible."""
    import torch

    prob = problems.struct_cube(3, quad=quad | kern)
    prob.Ag = np.random.default_rng(5).normal(size=prob.Ag.shape)
    say("struct assembly", prob.nNo, prob.mesh.nEl, "quad", quad, "kernel", hex(kern))
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    say("struct relerr", relerr(Vg, Vo), relerr(Rg, Ro))
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12
    st.assemble()
    torch.cuda.synchronize()
    assert np.array_equal(st.Val.cpu().numpy(), Vg)
    assert np.array_equal(st.R.cpu().numpy(), Rg)


@pytest.mark.parametrize("which", ["fluid", "lelas", "struct"])
def test_assembly_other_scatter_path(gpu, orc, monkeypatch, which):
    """Each element kernel also runs with the other DOASSEM strategy (tet4:
    the coloured read-modify-write; tet10: the gather); FP64 results within
    1e-12 of the oracle."""
    monkeypatch.setenv("SVFSI_ASM_GATHER", "1" if which == "struct" else "0")
    prob = {"fluid": lambda: problems.fluid_cylinder(h=0.4), "lelas": lambda: problems.lelas_cube(6),
            "struct": lambda: problems.struct_cube(3, quad=1 | SVFSI_STRUCT_MFMA64)}[which]()
    prob.Ag = np.random.default_rng(5).normal(size=prob.Ag.shape)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    say(which, "other scatter path relerr", relerr(Vg, Vo), relerr(Rg, Ro))
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12


@pytest.mark.parametrize("quad", [0, 1])
def test_struct_assembly_mfma32(gpu, orc, quad):
    """B^T Dm B on FP32 matrix cores with FP32 accumulation: the assembled
    matrix stays within 1e-6 of the FP64 oracle (max entry; FP32 rounding of
    ~6 nG products per entry), the residual is unchanged (FP64), and the
    sparsity and the FP64 kernel's bits are untouched."""
    prob = problems.struct_cube(3, quad=quad | SVFSI_STRUCT_MFMA32)
    prob.Ag = np.random.default_rng(5).normal(size=prob.Ag.shape)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    e = relerr(Vg, Vo)
    say("struct mfma32 relerr", e, relerr(Rg, Ro))
    assert 1e-12 < e < 1e-6
    assert relerr(Rg, Ro) < 1e-12


def test_fluid_assembly_parity_benchmark_mesh(gpu, orc):
    """The ~200k-tet benchmark cylinder, slivers included: FP64 assembly
    within 1e-12 of the oracle (max entry) and bitwise reproducible."""
    prob = problems.fluid_cylinder(h=0.11)
    say("fluid assembly (benchmark mesh)", prob.nNo, prob.mesh.nEl)
    st, Vg, Rg, Vo, Ro = _assembled(prob, orc)
    say("relerr", relerr(Vg, Vo), relerr(Rg, Ro))
    assert relerr(Vg, Vo) < 1e-12
    assert relerr(Rg, Ro) < 1e-12


def _solve_pair(prob, orc, mixed=False):
    st = NewtonStep(prob)
    st.assemble()
    import torch

    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy()
    xo, Vs, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), prob.ls, R0, V0)
    return st, ls, xg, xo, so, V0, R0, Vs


def test_cg_solve_parity(gpu, orc):
    prob = problems.lelas_cube(6)
    say("cg solve", prob.nNo)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc)
    assert ls.RI.suc == 1 and so.suc == 1
    assert abs(ls.RI.itr - so.itr) <= 2, (ls.RI.itr, so.itr)
    assert relerr(xg, xo) < 1e-7
    assert relerr(st.Val.cpu().numpy(), Vs) < 1e-13  # Val scaled in place like the reference
    assert abs(ls.RI.iNorm - so.iNorm) <= 1e-12 * so.iNorm


def test_gmres_solve_parity(gpu, orc):
    prob = problems.fluid_cylinder(h=0.4)
    say("gmres solve", prob.nNo)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc)
    say("itr gpu", ls.RI.itr, "cpu", so.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and so.suc == 1
    # restarted GMRES over ~1500 iterations: rounding may shift the exit
    # by a few Arnoldi steps
    assert abs(ls.RI.itr - so.itr) <= max(3, so.itr // 50), (ls.RI.itr, so.itr)
    # both meet the scaled FP64 criterion on the oracle's operator
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(Rs)
    assert relerr(xg, xo) < 1e-4
    assert abs(ls.RI.iNorm - so.iNorm) <= 1e-12 * so.iNorm


@pytest.mark.parametrize("cgs", ["2", "3"])
@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_mixed_outer_gram_schmidt(gpu, orc, monkeypatch, cgs, which):
    """The outer loop's Gram-Schmidt variants, two passes (2) and one pass
    with the explicit norm (3, the default): both stop on the residual
    estimate, and the FP64 residual recomputed on the CPU meets it."""
    monkeypatch.setenv("SVFSI_OUTER_CGS", cgs)
    prob = problems.fluid_cylinder(h=0.4) if which == "fluid" else problems.lelas_cube(6)
    say("outer gram-schmidt", cgs, which, prob.nNo)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=True)
    say("  outer", ls.RI.itr, "spmv", ls.RI.cM, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
    assert relerr(xg, xo) < 1e-4


@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_outer_norm_scale_fused(gpu, monkeypatch, which):
    """The explicit-norm pass's end in one kernel (k_gm_normscale, the
    default) against k_gm_givens pass 3 then k_gm_scale: bit for bit."""
    import torch

    prob = problems.fluid_cylinder(h=0.4) if which == "fluid" else problems.lelas_cube(6)
    st = NewtonStep(prob)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("SVFSI_GM_NORMSCALE", fused)
        st.assemble()
        ls = st.solve(mixed=True)
        torch.cuda.synchronize()
        out[fused] = (ls.RI.itr, ls.RI.cM, ls.RI.suc, st.R.cpu().numpy().copy())
        say("normscale", fused, which, "outer", ls.RI.itr)
    assert out["1"][:3] == out["0"][:3] and out["1"][2] == 1
    assert np.array_equal(out["1"][3], out["0"][3])


@pytest.mark.parametrize("mixed", [False, True])
def test_gmres_folded_partials(gpu, orc, monkeypatch, mixed):
    """The two-stage sum of the Arnoldi multi-dot partials (k_gm_fold, the
    default above 1024 row chunks, i.e. for large systems) forced on a small
    one: GMRESV and the mixed solve still match the oracle."""
    monkeypatch.setenv("SVFSI_FOLD_CHUNKS", "4")
    prob = problems.fluid_cylinder(h=0.4)
    say("gmres folded partials", prob.nNo, "mixed", mixed)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=mixed)
    say("itr gpu", ls.RI.itr, "cpu", so.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and so.suc == 1
    if not mixed:
        assert abs(ls.RI.itr - so.itr) <= max(3, so.itr // 50), (ls.RI.itr, so.itr)
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(Rs)
    assert relerr(xg, xo) < 1e-4


@pytest.mark.parametrize("mixed", [False, True])
def test_struct_solve(gpu, orc, mixed):
    prob = problems.struct_cube(4, quad=1)
    say("struct solve", prob.nNo, "mixed", mixed)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=mixed)
    say("itr gpu", ls.RI.itr, "cpu", so.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and so.suc == 1
    if not mixed:
        assert abs(ls.RI.itr - so.itr) <= max(3, so.itr // 50), (ls.RI.itr, so.itr)
        assert abs(ls.RI.iNorm - so.iNorm) <= 1e-12 * so.iNorm
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 2e-8 * np.linalg.norm(Rs)
    assert relerr(xg, xo) < 1e-4


def test_host_entry_point(gpu, orc):
    prob = problems.lelas_cube(4)
    Vo, Ro = orc.assemble(prob, *[a for a in orc.lhsa(prob.nNo, prob.mesh.ien)])
    rp, cp = fsils.lhsa(prob.nNo, prob.mesh.ien)
    lhs = fsils.fsils_lhs_create(prob.nNo, len(cp), rp, cp, nFaces=1)
    nodes, fdof = prob.dir_faces[0]
    fsils.fsils_bc_create(lhs, 1, len(nodes), fdof, fsils.BC_TYPE_Dir, nodes + 1)
    ls = fsils.fsils_ls_create(fsils.LS_TYPE_CG, relTol=1e-8, absTol=1e-30, maxItr=5000)
    Ri, Val = Ro.copy(), Vo.copy()
    fsils.fsils_solve(lhs, ls, 3, Ri, Val)
    xo, Vs, so = orc.fsils_solve(rp - 1, cp - 1, 3, prob.dir_mask(), dict(LS_type=798, relTol=1e-8, absTol=1e-30, mItr=5000), Ro, Vo)
    assert ls.RI.suc == 1
    assert relerr(Ri, xo) < 1e-7
    assert relerr(Val, Vs) < 1e-13


def test_mixed_reaches_fp64_tolerance(gpu, orc):
    prob = problems.fluid_cylinder(h=0.4)
    say("mixed solve", prob.nNo)
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=True)
    say("mixed outer", ls.RI.itr, "inner spmv", ls.RI.cM, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1
    # FP64 residual of the scaled system, recomputed on the CPU
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
    assert relerr(xg, xo) < 1e-4


@pytest.mark.parametrize("bj", ["0", "1"])
@pytest.mark.parametrize("which", ["fluid", "lelas", "struct"])
def test_polynomial_inner_preconditioner(gpu, orc, monkeypatch, bj, which):
    """Mixed solve with the GMRES(30) residual polynomial as the inner
    preconditioner (with and without block-Jacobi; dof 4 and dof 3): reaches
    the FP64 tolerance, and repeats bit for bit."""
    import torch

    monkeypatch.setenv("SVFSI_INNER_POLY", "1")
    monkeypatch.setenv("SVFSI_INNER_BJ", bj)
    prob = {"fluid": lambda: problems.fluid_cylinder(h=0.4), "lelas": lambda: problems.lelas_cube(6),
            "struct": lambda: problems.struct_cube(3, quad=1)}[which]()
    st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=True)
    say("poly bj", bj, "outer", ls.RI.itr, "SpMVs", ls.RI.cM)
    assert ls.RI.suc == 1
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
    assert relerr(xg, xo) < 1e-4
    st.assemble()
    ls2 = st.solve(mixed=True)
    torch.cuda.synchronize()
    assert (ls2.RI.itr, ls2.RI.cM) == (ls.RI.itr, ls.RI.cM)
    assert np.array_equal(st.R.cpu().numpy(), xg)


@pytest.mark.parametrize("poly", ["0", "1"])
@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_nested_solver_with_fp64_inner(gpu, orc, monkeypatch, poly, which):
    """SVFSI_INNER_FP64=1: the mixed entry point runs the same nested solver
    (FGMRES outer, polynomial or inner-GMRES preconditioner, block Jacobi as
    for FP32) with the inner matrix, vectors and products in FP64.  It
    reaches the FP64 tolerance on the oracle's operator, and takes no more
    outer steps than the FP32-inner solve (rounding of a FP32 inner
    operator can only weaken the preconditioner)."""
    monkeypatch.setenv("SVFSI_INNER_POLY", poly)
    prob = {"fluid": lambda: problems.fluid_cylinder(h=0.4), "lelas": lambda: problems.lelas_cube(6)}[which]()
    out = {}
    for inner64 in ("0", "1"):
        monkeypatch.setenv("SVFSI_INNER_FP64", inner64)
        st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=True)
        mask = prob.dir_mask()
        _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
        y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
        r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
        say("inner fp64", inner64, which, "poly", poly, "outer", ls.RI.itr, "SpMVs", ls.RI.cM,
            "x relerr", relerr(xg, xo))
        assert ls.RI.suc == 1
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
        assert relerr(xg, xo) < 1e-4
        out[inner64] = ls.RI.itr
    assert out["1"] <= out["0"] + 2, out


def test_inner_block_jacobi_reduces_inner_iterations(gpu, orc, monkeypatch):
    """The FP32 inner cycle of the mixed solve is left-preconditioned by the
    inverse 4x4 node blocks.  Both variants must reach the FP64 tolerance;
    the block-Jacobi one with fewer inner SpMVs."""
    prob = problems.fluid_cylinder(h=0.4)
    monkeypatch.setenv("SVFSI_INNER_POLY", "0")
    res = {}
    for bj in ("0", "1"):
        monkeypatch.setenv("SVFSI_INNER_BJ", bj)
        st, ls, xg, xo, so, V0, R0, Vs = _solve_pair(prob, orc, mixed=True)
        mask = prob.dir_mask()
        _, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, mask)
        y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
        r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, prob.dof, Vs, y)
        assert ls.RI.suc == 1
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
        res[bj] = ls.RI.cM
        say("block-Jacobi", bj, "outer", ls.RI.itr, "SpMVs", ls.RI.cM)
    assert res["1"] < res["0"], res


@pytest.mark.parametrize("mixed,cg,poly", [(True, False, "0"), (True, False, "1"), (True, True, "1"),
                                           (False, False, "1"), (False, True, "1")])
def test_graph_replays_match_direct_launches(gpu, monkeypatch, mixed, cg, poly):
    """The solvers replay fixed kernel sequences as HIP graphs (one per outer
    step of the mixed solve with the inner GMRES cycle, one per chunk of
    Arnoldi steps of the FP64 GMRES, one per chunk of CG iterations; the
    polynomial steps are launched directly either way).  Several replays,
    with eager work and synchronisation between them, must reproduce the
    directly launched solve bit for bit."""
    import torch

    prob = problems.lelas_cube(8) if cg else problems.fluid_cylinder(h=0.5)
    st = NewtonStep(prob)
    monkeypatch.setenv("SVFSI_INNER_POLY", poly)
    monkeypatch.setenv("SVFSI_GRAPHS", "0")
    st.assemble()
    ls0 = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    x0 = st.R.cpu().numpy().copy()
    monkeypatch.setenv("SVFSI_GRAPHS", "1")
    for rep in range(3):
        st.assemble()
        ls = st.solve(mixed=mixed)
        torch.cuda.synchronize()
        say("graph replay", mixed, rep, ls.RI.itr, ls.RI.cM)
        assert (ls.RI.itr, ls.RI.cM, ls.RI.suc) == (ls0.RI.itr, ls0.RI.cM, ls0.RI.suc)
        assert np.array_equal(st.R.cpu().numpy(), x0)
        assert float((torch.arange(4096, device=gpu, dtype=torch.float64) * 0.5).sum().item()) == 4096 * 4095 / 4


def test_lelas_dkde(gpu, orc):
    import torch

    prob = problems.lelas_cube(4)
    st = NewtonStep(prob)
    rng = np.random.default_rng(9)
    lam = rng.normal(size=(prob.nNo, 3))
    u = rng.normal(size=(prob.nNo, 3))
    g = torch.zeros(prob.mesh.nEl, dtype=torch.float64, device=gpu)
    st.dmesh.lelas_dkde_dot(prob.eq, T(lam, gpu), T(u, gpu), g)
    torch.cuda.synchronize()
    go = orc.lelas_dkde_dot(prob, lam, u)
    assert relerr(g.cpu().numpy(), go) < 1e-12
