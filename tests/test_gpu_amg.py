"""The multigrid inner preconditioner of the mixed solve (SVFSI_INNER_AMG=1,
amg.hip; no reference counterpart): the FGMRES outer loop with one FP32
V-cycle per step reaches the reference's FP64 criterion on the oracle's
scaled operator (residual recomputed on the CPU), matches the oracle's
solution, repeats bit for bit, and builds a coarsening hierarchy."""
import ctypes as C
import sys

import numpy as np
import pytest

from svfsi_amd import problems
from svfsi_amd._lib import lib
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("which", ["lelas", "struct", "fluid"])
def test_amg_inner_reaches_fp64_tolerance(gpu, orc, monkeypatch, which):
    import torch

    monkeypatch.setenv("SVFSI_INNER_AMG", "1")
    prob = {"fluid": lambda: problems.fluid_cylinder(h=0.4), "lelas": lambda: problems.lelas_cube(8),
            "struct": lambda: problems.struct_cube(4, quad=1)}[which]()
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    say("amg", which, "nNo", prob.nNo)
    ls = st.solve(mixed=True)
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy().copy()
    lv = (C.c_int * 16)()
    nl = lib().svfsi_amg_levels(st.lhs.handle, lv, 16)
    sizes = list(lv)[:nl]
    say("levels", sizes, "outer", ls.RI.itr, "SpMVs", ls.RI.cM, "suc", ls.RI.suc)
    assert nl >= 2 and sizes[0] == prob.nNo and all(b < a for a, b in zip(sizes, sizes[1:]))
    assert ls.RI.suc == 1
    rp, cp = st.rowPtr - 1, st.colPtr - 1
    mask = prob.dir_mask()
    _, Rs, W = orc.precond_diag(rp, cp, prob.dof, V0, R0, mask)
    xo, Vs, so = orc.fsils_solve(rp, cp, prob.dof, mask, prob.ls, R0, V0)
    y = np.where(W != 0, xg / np.where(W != 0, W, 1), 0.0)
    r = Rs - orc.spmv(rp, cp, prob.dof, Vs, y)
    say("scaled residual", np.linalg.norm(r) / np.linalg.norm(Rs), "x relerr", relerr(xg, xo))
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Rs) * 1.0001
    assert relerr(xg, xo) < 1e-4
    # bitwise repeat (the hierarchy's pattern part is reused)
    st.assemble()
    ls2 = st.solve(mixed=True)
    torch.cuda.synchronize()
    assert (ls2.RI.itr, ls2.RI.cM) == (ls.RI.itr, ls.RI.cM)
    assert np.array_equal(st.R.cpu().numpy(), xg)


@pytest.mark.parametrize("which", ["lelas", "fluid"])
def test_amg_fused_kernels_match_separate_ones(gpu, monkeypatch, which):
    """The V-cycle's residual and Chebyshev steps fused with their products
    on the row tiles (the default) give the separate kernels' results bit
    for bit (SVFSI_AMG_FUSED=0): the same row sums, the same arithmetic."""
    import torch

    monkeypatch.setenv("SVFSI_INNER_AMG", "1")
    prob = {"fluid": lambda: problems.fluid_cylinder(h=0.4), "lelas": lambda: problems.lelas_cube(8)}[which]()
    st = NewtonStep(prob)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("SVFSI_AMG_FUSED", fused)
        st.assemble()
        ls = st.solve(mixed=True)
        torch.cuda.synchronize()
        out[fused] = (ls.RI.itr, ls.RI.cM, ls.RI.suc, st.R.cpu().numpy().copy())
        say("amg fused", fused, which, "outer", ls.RI.itr, "SpMVs", ls.RI.cM)
    assert out["1"][:3] == out["0"][:3]
    assert out["1"][2] == 1
    assert np.array_equal(out["1"][3], out["0"][3])
