"""The allocation-free CG solve (fsils_gpu_solve_cg_noalloc_dev) through the
C ABI against the existing entry point it is a variant of
(fsils_gpu_solve_dev with LS_TYPE_CG and PRECOND_FSILS): two calls in a
row, the second on a smaller system, each compared bit for bit with the
existing entry point (same kernels in the same order) and with the oracle;
and a system beyond the fixed scratch size is refused before any GPU work."""
import ctypes as C
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, problems
from svfsi_amd._lib import SVFSI_ERR_ARG, lib
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc_solve(st: NewtonStep, ls, Ri, Val) -> int:
    return lib().fsils_gpu_solve_cg_noalloc_dev(st.lhs.handle, C.byref(ls), st.prob.dof, fsils.dev_ptr(Ri),
                                                fsils.dev_ptr(Val), None, fsils.cur_stream())


def test_noalloc_cg_matches_the_allocating_entry_point(gpu, orc):
    import torch

    for n in (8, 5):  # the second call on a smaller system
        prob = problems.lelas_cube(n)
        prob.ls = dict(prob.ls, relTol=1e-10, mItr=5000)
        st = NewtonStep(prob)
        st.assemble()
        torch.cuda.synchronize()
        V0, R0 = st.Val.clone(), st.R.clone()
        say("noalloc CG: lelas_cube", n, "unknowns", prob.nNo * prob.dof)
        ls_ref = st.make_ls()
        fsils.fsils_solve(st.lhs, ls_ref, prob.dof, st.R, st.Val)
        torch.cuda.synchronize()
        x_ref = st.R.cpu().numpy().copy()
        ls_na = st.make_ls()
        Ri, Val = R0.clone(), V0.clone()
        rc = noalloc_solve(st, ls_na, Ri, Val)
        torch.cuda.synchronize()
        assert rc == 0, lib().svfsi_last_error()
        x_na = Ri.cpu().numpy()
        say("itr", ls_na.RI.itr, ls_ref.RI.itr, "fNorm", ls_na.RI.fNorm, ls_ref.RI.fNorm)
        assert (ls_na.RI.suc, ls_na.RI.itr, ls_na.RI.cM) == (ls_ref.RI.suc, ls_ref.RI.itr, ls_ref.RI.cM)
        assert ls_na.RI.iNorm == ls_ref.RI.iNorm and ls_na.RI.fNorm == ls_ref.RI.fNorm
        assert np.array_equal(x_na, x_ref)
        # Val is left scaled in place, as by the existing entry point
        assert torch.equal(Val, st.Val)
        xo, _, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(),
                                    prob.ls, R0.cpu().numpy(), V0.cpu().numpy())
        assert ls_na.RI.suc == 1 and abs(ls_na.RI.itr - so.itr) <= 2
        assert np.abs(x_na - xo).max() <= 1e-8 * np.abs(xo).max()


def test_noalloc_cg_refuses_systems_beyond_its_scratch(gpu):
    import torch

    cap = lib().svfsi_noalloc_scratch_bytes()
    nNo = cap // (5 * 8 * 3) + 1  # dof 3: five vectors of nNo * 3 doubles exceed the scratch
    rp = np.arange(1, nNo + 2, dtype=np.int32)
    cp = np.arange(1, nNo + 1, dtype=np.int32)
    lhs = fsils.fsils_lhs_create(nNo, nNo, rp, cp)
    ls = fsils.fsils_ls_create(fsils.LS_TYPE_CG)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")  # never touched: refused first
    rc = lib().fsils_gpu_solve_cg_noalloc_dev(lhs.handle, C.byref(ls), 3, fsils.dev_ptr(dummy),
                                              fsils.dev_ptr(dummy), None, fsils.cur_stream())
    msg = lib().svfsi_last_error().decode()
    say("refused:", msg)
    assert rc == SVFSI_ERR_ARG and "scratch" in msg
    ls2 = fsils.fsils_ls_create(fsils.LS_TYPE_GMRES)
    rc = lib().fsils_gpu_solve_cg_noalloc_dev(lhs.handle, C.byref(ls2), 3, fsils.dev_ptr(dummy),
                                              fsils.dev_ptr(dummy), None, fsils.cur_stream())
    assert rc != 0 and "CGRADV only" in lib().svfsi_last_error().decode()
