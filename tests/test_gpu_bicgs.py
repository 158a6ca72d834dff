"""GPU BICGSV (svFSILS/BICGS.f:115-180) through FSILS_SOLVE with
LS_TYPE_BICGS against the oracle on the same assembled systems: iteration
count, success flag and solution.  The mixed entry point runs its flexible
GMRES for every LS_type but NS: with LS_TYPE_BICGS it makes exactly the
iterations it makes with LS_TYPE_GMRES."""
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, problems
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_bicgs_parity(gpu, orc, which):
    import torch

    prob = problems.fluid_cylinder(h=0.4) if which == "fluid" else problems.lelas_cube(5)
    settings = dict(LS_type=fsils.LS_TYPE_BICGS, relTol=1e-8, absTol=1e-30, mItr=5000)
    prob.ls = dict(settings)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve()
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy()
    xo, _, lo = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), settings, R0, V0)
    say("bicgs", which, "itr", ls.RI.itr, lo.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and lo.suc == 1
    # BiCGStab's residual is not monotone: summation-order differences move
    # the iteration where it first drops below the tolerance by a few steps
    # (GMRES and CG stay within 2 %)
    assert abs(ls.RI.itr - lo.itr) <= max(5, lo.itr // 10)
    assert ls.RI.fNorm < 1e-8 * ls.RI.iNorm and abs(ls.RI.iNorm - lo.iNorm) <= 1e-12 * lo.iNorm
    assert relerr(xg, xo) < 1e-5


def test_mixed_entry_point_runs_fgmres_for_bicgs(gpu):
    import torch

    out = {}
    for LS_type in (fsils.LS_TYPE_GMRES, fsils.LS_TYPE_BICGS):
        prob = problems.lelas_cube(5)
        prob.ls = dict(LS_type=LS_type, relTol=1e-8, absTol=1e-30, mItr=50, sD=30)
        st = NewtonStep(prob)
        st.assemble()
        ls = st.solve(mixed=True)
        torch.cuda.synchronize()
        out[LS_type] = (ls.RI.suc, ls.RI.itr, ls.RI.cM, st.R.cpu().numpy().copy())
    say("mixed gmres vs bicgs", out[fsils.LS_TYPE_GMRES][:3], out[fsils.LS_TYPE_BICGS][:3])
    assert out[fsils.LS_TYPE_GMRES][:3] == out[fsils.LS_TYPE_BICGS][:3]
    assert np.array_equal(out[fsils.LS_TYPE_GMRES][3], out[fsils.LS_TYPE_BICGS][3])
