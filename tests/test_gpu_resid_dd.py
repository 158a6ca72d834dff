"""svfsi_resid_dd_dev: b - K x and b - K^T x with double-double row sums
(Ogita-Rump-Oishi Dot2), against the exact residual computed in rational
arithmetic (fractions.Fraction) and rounded once.  The right-hand side is
K x to within 1e-10 relative, so the residual is ~1e-10 of the products it
is formed from: the plain FP64 product loses most of its digits there, the
compensated one must stay within Dot2's bound
|r - r_exact| <= eps |r_exact| + (n eps)^2 sum |a x|."""
import sys
from fractions import Fraction

import numpy as np
import pytest

from svfsi_amd import fsils, problems
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def exact_resid(Val, rp, cp, dof, x, b, transpose):
    nNo = len(rp) - 1
    K = {}
    for i in range(nNo):
        for k in range(rp[i], rp[i + 1]):
            blk = Val[k].reshape(dof, dof)
            K[(i, cp[k])] = blk
    out = np.zeros_like(b)
    bound = np.zeros_like(b)
    nterm = 0
    for i in range(nNo):
        for p in range(dof):
            acc = Fraction(b[i, p])
            mag = 0.0
            terms = 0
            for k in range(rp[i], rp[i + 1]):
                c = cp[k]
                for q in range(dof):
                    a = K[(c, i)][q, p] if transpose else K[(i, c)][p, q]
                    acc -= Fraction(a) * Fraction(x[c, q])
                    mag += abs(a * x[c, q])
                    terms += 1
            out[i, p] = float(acc)
            bound[i, p] = EPS * abs(out[i, p]) + ((terms + 1) * EPS) ** 2 * mag * 2.0
            nterm = max(nterm, terms)
    return out, bound


@pytest.mark.parametrize("transpose", [0, 1])
def test_resid_dd_matches_the_exact_residual(gpu, transpose):
    import torch

    prob = problems.lelas_cube(3)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    Val = st.Val.cpu().numpy().copy()
    rp, cp = st.rowPtr - 1, st.colPtr - 1
    rng = np.random.default_rng(7 + transpose)
    x = rng.normal(size=(prob.nNo, 3))
    # b = K x (exact, rounded) times (1 + 1e-10 noise)
    b0, _ = exact_resid(Val, rp, cp, 3, x, np.zeros((prob.nNo, 3)), transpose)
    b = -b0 * (1.0 + 1e-10 * rng.normal(size=b0.shape))
    r_exact, bound = exact_resid(Val, rp, cp, 3, x, b, transpose)
    dV, dx, db = (torch.from_numpy(a).to(gpu) for a in (Val, x, b))
    r = torch.empty_like(db)
    fsils.call("svfsi_resid_dd_dev", st.lhs.handle, 3, fsils.dev_ptr(dV), fsils.dev_ptr(dx), fsils.dev_ptr(db),
               fsils.dev_ptr(r), transpose, fsils.cur_stream())
    # the plain FP64 product for comparison
    Kx = torch.empty_like(db)
    fsils.call("svfsi_spmv_t_dev" if transpose else "svfsi_spmv_dev", st.lhs.handle, 3, fsils.dev_ptr(dV),
               fsils.dev_ptr(dx), fsils.dev_ptr(Kx), fsils.cur_stream())
    torch.cuda.synchronize()
    rd = r.cpu().numpy()
    rn = b - Kx.cpu().numpy()
    err_dd = np.abs(rd - r_exact)
    err_plain = np.abs(rn - r_exact).max() / np.abs(r_exact).max()
    print("transpose", transpose, "dd rel", err_dd.max() / np.abs(r_exact).max(), "plain rel", err_plain,
          flush=True)
    sys.stdout.flush()
    assert np.all(err_dd <= bound), (err_dd / bound).max()
    assert err_plain > 1e3 * (err_dd.max() / np.abs(r_exact).max() + 1e-300)  # the case needs the compensation
    # aliasing dR = dB is allowed
    fsils.call("svfsi_resid_dd_dev", st.lhs.handle, 3, fsils.dev_ptr(dV), fsils.dev_ptr(dx), fsils.dev_ptr(db),
               fsils.dev_ptr(db), transpose, fsils.cur_stream())
    torch.cuda.synchronize()
    assert np.array_equal(db.cpu().numpy(), rd)
