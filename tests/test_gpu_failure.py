"""Failure detection of the FP64 solves on the GPU, forced from the host
through the inputs (no kernel is made to fault): the outputs the reference
reports in these cases (svFSILS/GMRES.f:273-431, CGRAD.f:184-242,
BICGS.f:115-180, SOLVE.f:51-143).

- Iteration cap reached: suc = 0, with the reference's itr and the final
  residual of the capped iteration, matched against the oracle run with the
  same cap.
- Zero right-hand side: GMRESV's early return below absTol
  (GMRES.f:297, 307-311) leaves suc false, with itr = 0, dB = 0, x = 0;
  CGRADV and BICGSV pass their first stop test (CGRAD.f:211-214,
  BICGS.f:146-150): suc = 1, itr = 0, dB = 0 (errO below EPSILON), x = 0.
- A NaN in the right-hand side: every solver terminates (its stop test never
  passes), reports suc = 0 and a NaN residual, and the host call returns.
"""
import math
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, problems
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(which, ls):
    import torch

    prob = problems.fluid_cylinder(h=0.4) if which == "fluid" else problems.lelas_cube(5)
    prob.ls = dict(prob.ls, **ls)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    return prob, st, st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()


@pytest.mark.parametrize("which,ls", [
    ("fluid", dict(LS_type=fsils.LS_TYPE_GMRES, relTol=1e-12, absTol=1e-30, mItr=2, sD=5)),
    ("lelas", dict(LS_type=fsils.LS_TYPE_CG, relTol=1e-12, absTol=1e-30, mItr=7, sD=0)),
])
def test_iteration_cap_reports_failure_like_the_oracle(gpu, orc, which, ls):
    import torch

    prob, st, V0, R0 = setup(which, ls)
    res = st.solve()
    torch.cuda.synchronize()
    x = st.R.cpu().numpy()
    xo, _, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), prob.ls, R0, V0)
    say(which, "gpu suc/itr/fNorm", res.RI.suc, res.RI.itr, res.RI.fNorm, "oracle", so.suc, so.itr, so.fNorm)
    assert res.RI.suc == 0 and so.suc == 0
    assert res.RI.itr == so.itr
    assert abs(res.RI.fNorm - so.fNorm) <= 1e-8 * so.fNorm
    assert res.RI.fNorm > ls["relTol"] * res.RI.iNorm
    assert np.abs(x - xo).max() <= 1e-8 * np.abs(xo).max()


@pytest.mark.parametrize("LS_type", [fsils.LS_TYPE_GMRES, fsils.LS_TYPE_CG, fsils.LS_TYPE_BICGS])
def test_zero_rhs_returns_at_once(gpu, orc, LS_type):
    import torch

    prob, st, V0, _ = setup("lelas", dict(LS_type=LS_type, relTol=1e-8, absTol=1e-10, mItr=50, sD=10))
    st.R.zero_()
    res = st.solve()
    torch.cuda.synchronize()
    _, _, so = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), prob.ls,
                               np.zeros((prob.nNo, prob.dof)), V0)
    say("zero rhs", LS_type, res.RI.suc, res.RI.itr, res.RI.iNorm, res.RI.dB, "oracle", so.suc, so.itr, so.dB)
    assert (res.RI.suc, res.RI.itr) == (so.suc, so.itr)
    assert res.RI.suc == (0 if LS_type == fsils.LS_TYPE_GMRES else 1)
    assert res.RI.itr == 0 and res.RI.iNorm == 0.0 and res.RI.dB == 0.0
    assert torch.count_nonzero(st.R).item() == 0


@pytest.mark.parametrize("LS_type,mixed", [(fsils.LS_TYPE_GMRES, False), (fsils.LS_TYPE_CG, False),
                                           (fsils.LS_TYPE_BICGS, False), (fsils.LS_TYPE_GMRES, True)])
def test_nan_rhs_terminates_with_failure(gpu, LS_type, mixed):
    import torch

    prob, st, _, _ = setup("lelas", dict(LS_type=LS_type, relTol=1e-8, absTol=1e-30, mItr=4, sD=10))
    # a free dof of an interior node
    free = np.nonzero(prob.dir_mask()[:, 0] == 1.0)[0]
    st.R[int(free[len(free) // 2]), 0] = float("nan")
    say("nan rhs", LS_type, "mixed", mixed)
    res = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    say("  suc", res.RI.suc, "itr", res.RI.itr, "iNorm", res.RI.iNorm, "fNorm", res.RI.fNorm)
    assert res.RI.suc == 0
    assert math.isnan(res.RI.iNorm) or math.isnan(res.RI.fNorm)
