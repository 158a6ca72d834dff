"""Writes the regression fixtures in tests/golden/ from the oracle.

The reference holds no golden vectors for this path, so these fixtures are
the oracle's own outputs on tiny seeded inputs, frozen so that a later change
of the oracle or of the HIP path shows up as a difference.  Each .npz holds
the inputs (mesh, state, parameters, pattern) and the expected outputs.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import oracle as orc  # noqa: E402
from svfsi_amd import mesh, problems  # noqa: E402


def eq_json(eq):
    return json.dumps({k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v) for k, v in eq.items()})


def save(name, prob, extra=None):
    m = prob.mesh
    rp, cp = orc.lhsa(m.nNo, m.ien)
    V, R = orc.assemble(prob, rp, cp)
    d = dict(x=m.x, ien=m.ien, Ag=prob.Ag, Yg=prob.Yg, Dg=prob.Dg, phys=prob.phys, dof=prob.dof, quad=prob.quad,
             eq=eq_json(prob.eq), rowPtr=rp, colPtr=cp, Val=V, R=R,
             prop_el=prob.prop_el if prob.prop_el is not None else np.zeros(0))
    if prob.trac_faces:
        (tf, te, th), = prob.trac_faces
        d.update(trac_fien=tf, trac_fel=te, trac_h=th)
    if extra:
        d.update(extra(prob, rp, cp, V, R))
    np.savez_compressed(HERE / f"{name}.npz", **d)
    print(name, m.nNo, m.nEl, len(cp))


def lelas():
    p = problems.lelas_cube(2)
    rng = np.random.default_rng(21)
    p.Dg = rng.normal(size=p.Dg.shape) * 1e-3
    p.Ag = rng.normal(size=p.Ag.shape)
    p.prop_el = 1e6 * np.exp(0.3 * rng.normal(size=p.mesh.nEl))

    def extra(prob, rp, cp, V, R):
        ls = dict(prob.ls, relTol=1e-10)
        x, Vs, so = orc.fsils_solve(rp, cp, 3, prob.dir_mask(), ls, R, V)
        lam = rng.normal(size=(prob.nNo, 3))
        return dict(mask=prob.dir_mask(), sol=x, itr=so.itr, lam=lam, g=orc.lelas_dkde_dot(prob, lam, x))

    save("lelas_cube2", p, extra)


def fluid():
    m = mesh.cube_tet4(1)
    rng = np.random.default_rng(22)
    eq = dict(problems.gen_alpha(0.2, False), dt=1e-3, rho=1.06, mu=0.04, bstab=0.2)
    Y = rng.uniform(-1, 1, (m.nNo, 4))
    A = rng.normal(size=(m.nNo, 4))
    p = problems.Problem(name="fluid_cube1", phys=problems.SVFSI_PHYS_FLUID, dof=4, mesh=m, eq=eq, Ag=A, Yg=Y,
                         Dg=np.zeros_like(Y), dir_faces=[], prop_el=0.04 * np.exp(0.1 * rng.normal(size=m.nEl)))

    def extra(prob, rp, cp, V, R):
        lam = rng.normal(size=(prob.nNo, 4))
        du = rng.normal(size=(prob.nNo, 4))
        return dict(lam=lam, du=du, g=orc.fluid_dmu_dot(prob, lam, du))

    save("fluid_cube1", p, extra)


def struct():
    for quad in (0, 1):
        p = problems.struct_cube(1, quad=quad, amp=0.02)
        p.Ag = np.random.default_rng(23).normal(size=p.Ag.shape)
        save(f"struct_cube1_q{quad}", p)


if __name__ == "__main__":
    lelas()
    fluid()
    struct()
