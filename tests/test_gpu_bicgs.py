"""GPU BICGSV (svFSILS/BICGS.f:115-180) through FSILS_SOLVE with
LS_TYPE_BICGS against the oracle on the same assembled systems: iteration
count, success flag and solution; and the mixed path with the same
LS_type (flexible GMRES, as for the other solver types)."""
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


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("which", ["fluid", "lelas"])
def test_bicgs_parity(gpu, orc, which, mixed):
    import torch

    prob = problems.fluid_cylinder(h=0.4) if which == "fluid" else problems.lelas_cube(5)
    settings = dict(LS_type=fsils.LS_TYPE_BICGS, relTol=1e-8, absTol=1e-30, mItr=5000)
    prob.ls = dict(settings)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    ls = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    xg = st.R.cpu().numpy()
    xo, _, lo = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), settings, R0, V0)
    say("bicgs", which, "mixed", mixed, "itr", ls.RI.itr, lo.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1 and lo.suc == 1
    if not mixed:
        # BiCGStab's residual is not monotone: summation-order differences
        # move the iteration where it first drops below the tolerance by a
        # few steps (GMRES and CG stay within 2 %)
        assert abs(ls.RI.itr - lo.itr) <= max(5, lo.itr // 10)
        assert ls.RI.fNorm < 1e-8 * ls.RI.iNorm and abs(ls.RI.iNorm - lo.iNorm) <= 1e-12 * lo.iNorm
    assert relerr(xg, xo) < 1e-5
