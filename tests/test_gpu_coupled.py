"""Coupled Neumann (resistance) faces on the GPU through the C ABI against
the oracle's FSILS_SOLVE with the same faces: GMRES (FP64 and mixed) and the
NS solver add ADDBCMUL's rank-one terms, CG ignores them (as CGRADV does),
PRECOND_RCS with coupled faces is refused.  Faces that share nodes take the
per-face launch path; a second solve with another res replays the captured
graphs and must see the new coefficient."""
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, problems
from svfsi_amd._lib import SvfsiError
from svfsi_amd.step import NewtonStep

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def outlet_faces(prob, split=False):
    ff, fe = prob.bflow_faces[0]
    nodes, val = problems.face_normal_integral(prob.mesh, ff, fe)
    if not split:
        return [(nodes, 3, val, 0.0)]
    # two faces sharing about a fifth of their nodes
    k = len(nodes)
    a, b = slice(0, (3 * k) // 5), slice((2 * k) // 5, k)
    return [(nodes[a], 3, val[a], 0.0), (nodes[b], 3, 0.5 * val[b], 0.0)]


def setup(orc, ls, split=False, which="fluid"):
    import torch

    if which == "fluid":
        prob = problems.fluid_cylinder(h=0.6)
        rng = np.random.default_rng(3)
        prob.Yg[:, :3] += 0.1 * rng.normal(size=(prob.nNo, 3))
        prob.res_faces = outlet_faces(prob, split)
    else:
        prob = problems.lelas_cube(4)
        ff, fe = prob.trac_faces[0][:2]
        nodes, val = problems.face_normal_integral(prob.mesh, ff, fe)
        prob.res_faces = [(nodes, 3, val, 0.0)]
    prob.ls = dict(prob.ls, **ls)
    st = NewtonStep(prob)
    st.assemble()
    torch.cuda.synchronize()
    V0, R0 = st.Val.cpu().numpy().copy(), st.R.cpu().numpy().copy()
    return prob, st, V0, R0


def res_values(orc, st, prob, V0, R0, target):
    """res per face with res |val W|^2 = target"""
    _, _, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, prob.dof, V0, R0, prob.dir_mask())
    out = []
    for nodes, fd, val, _ in prob.res_faces:
        vm = val * W[nodes, :fd]
        out.append(target / float((vm * vm).sum()))
    return out


def solve_gpu(st, prob, V0, R0, res, mixed=False):
    import torch

    st.Val.copy_(torch.from_numpy(V0))
    st.R.copy_(torch.from_numpy(R0))
    st.res = np.array([0.0] * len(prob.dir_faces) + list(res))
    ls = st.solve(mixed=mixed)
    torch.cuda.synchronize()
    return st.R.cpu().numpy(), ls


def solve_cpu(orc, st, prob, V0, R0, res, settings):
    faces = [(n, fd, v, r) for (n, fd, v, _), r in zip(prob.res_faces, res)]
    return orc.fsils_solve_cpl(st.rowPtr - 1, st.colPtr - 1, prob.dof, prob.dir_mask(), settings, R0, V0, faces)


GM = dict(LS_type=fsils.LS_TYPE_GMRES, relTol=1e-8, absTol=1e-30, mItr=1000, sD=50)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("target", [1.0, 100.0])
def test_gmres_resistance_parity(gpu, orc, target, split):
    prob, st, V0, R0 = setup(orc, GM, split)
    res = res_values(orc, st, prob, V0, R0, target)
    xg, ls = solve_gpu(st, prob, V0, R0, res)
    xo, _, lo = solve_cpu(orc, st, prob, V0, R0, res, GM)
    x0, _, l0 = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, 4, prob.dir_mask(), GM, R0, V0)
    say("gmres coupled target", target, "split", split, "itr", ls.RI.itr, lo.RI.itr, "(uncoupled", l0.itr,
        ") x relerr", relerr(xg, xo), "vs uncoupled", relerr(xg, x0))
    assert ls.RI.suc == lo.RI.suc == 1
    assert abs(ls.RI.itr - lo.RI.itr) <= max(3, lo.RI.itr // 50)
    assert relerr(xg, xo) < 1e-6
    assert relerr(xg, x0) > 1e-4  # the coupled term changed the answer


def test_gmres_resistance_graph_replay_sees_new_res(gpu, orc):
    """Same faces, two solves with different res: the second replays the
    captured Arnoldi graphs, whose coupled-face launches read res from
    device memory."""
    prob, st, V0, R0 = setup(orc, GM)
    for target in (1.0, 30.0):
        res = res_values(orc, st, prob, V0, R0, target)
        xg, ls = solve_gpu(st, prob, V0, R0, res)
        xo, _, lo = solve_cpu(orc, st, prob, V0, R0, res, GM)
        say("replay target", target, "itr", ls.RI.itr, lo.RI.itr, "x relerr", relerr(xg, xo))
        assert ls.RI.suc == 1 and relerr(xg, xo) < 1e-6


def test_mixed_resistance(gpu, orc):
    prob, st, V0, R0 = setup(orc, GM)
    res = res_values(orc, st, prob, V0, R0, 100.0)
    xg, ls = solve_gpu(st, prob, V0, R0, res, mixed=True)
    xo, _, lo = solve_cpu(orc, st, prob, V0, R0, res, GM)
    say("mixed coupled itr", ls.RI.itr, "fp64 cpu", lo.RI.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == 1
    assert relerr(xg, xo) < 1e-5


@pytest.mark.parametrize("split", [False, True])
def test_ns_resistance_parity(gpu, orc, split):
    # relTol 1e-5: the NS stopping test resolves iNorm^2 - xB.B only to
    # about 1e-12 iNorm^2
    settings = dict(LS_type=fsils.LS_TYPE_NS, relTol=1e-5, absTol=1e-30, mItr=20, sD=50,
                    GM=dict(relTol=1e-4, mItr=4), CG=dict(relTol=1e-4))
    prob, st, V0, R0 = setup(orc, settings, split)
    res = res_values(orc, st, prob, V0, R0, 10.0)
    xg, ls = solve_gpu(st, prob, V0, R0, res)
    xo, _, lo = solve_cpu(orc, st, prob, V0, R0, res, settings)
    say("ns coupled split", split, "RI.itr", ls.RI.itr, lo.RI.itr, "GM", ls.GM.itr, lo.GM.itr, "CG", ls.CG.itr,
        lo.CG.itr, "x relerr", relerr(xg, xo))
    assert ls.RI.suc == lo.RI.suc == 1
    assert ls.RI.itr == lo.RI.itr
    assert abs(ls.GM.itr - lo.GM.itr) <= max(2, lo.GM.itr // 20)
    assert abs(ls.CG.itr - lo.CG.itr) <= max(2, lo.CG.itr // 20)
    assert relerr(xg, xo) < 1e-6


def test_cg_ignores_resistance(gpu, orc):
    settings = dict(LS_type=fsils.LS_TYPE_CG, relTol=1e-8, absTol=1e-30, mItr=1000)
    prob, st, V0, R0 = setup(orc, settings, which="lelas")
    xg, ls = solve_gpu(st, prob, V0, R0, [1e3])
    x0, _, l0 = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, 3, prob.dir_mask(), settings, R0, V0)
    say("cg with a resistance face: itr", ls.RI.itr, l0.itr, "x relerr vs uncoupled", relerr(xg, x0))
    assert ls.RI.suc == 1 and abs(ls.RI.itr - l0.itr) <= 2
    assert relerr(xg, x0) < 1e-6


def test_rcs_with_resistance_refused(gpu, orc):
    prob, st, V0, R0 = setup(orc, dict(GM, prec=fsils.PRECOND_RCS))
    with pytest.raises(SvfsiError, match="PRECOND_FSILS"):
        solve_gpu(st, prob, V0, R0, [1.0])


def test_neu_face_without_res_refused(gpu, orc):
    prob, st, V0, R0 = setup(orc, GM)
    st.res = None
    with pytest.raises(SvfsiError, match="res is required"):
        st.solve()
