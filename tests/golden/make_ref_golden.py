"""Writes the reference fixtures tests/golden/ref_*.npz from the reference's
own Fortran (svFSI CONSTRUCT_FLUID / CONSTRUCT_LELAS / CONSTRUCT_dSOLID with
LHSA and DOASSEM, and svFSILS FSILS_SOLVE), built from the reference sources
by oracle/ref.mk and driven by oracle/ref_harness.f90.

Each fixture is one small seeded problem at a general state: a cube mesh
with jittered interior nodes, random nodal acceleration, velocity, pressure
and displacement, and parameters chosen so that every term of the element
routine is active (for FLUID3D the fine-scale velocity u' and the
discontinuity-capturing tau_B).  It holds the inputs (mesh, state,
parameters) and the reference's outputs: the pattern (LHSA), Val and R
(element loop + DOASSEM) and, for the solve fixtures, the reference's
increment and its solver record (iteration counts, norms).  Only this
script runs the reference; the tests read the committed .npz files.

    make -f oracle/ref.mk -j8 && python tests/golden/make_ref_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import refharness as rh  # noqa: E402
from svfsi_amd import mesh, problems  # noqa: E402
from svfsi_amd._lib import LS_TYPE_BICGS, LS_TYPE_CG, LS_TYPE_GMRES, LS_TYPE_NS  # noqa: E402


def eq_json(eq):
    return json.dumps({k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v) for k, v in eq.items()})


def jitter(m: mesh.TetMesh, frac: float, seed: int) -> mesh.TetMesh:
    """Moves every node not on the cube's boundary by up to frac of the
    lattice spacing in each direction (tet4 and tet10 corner lattice; a tet10
    mid-edge node moves with the mean of its edge's corners, so the edges
    stay straight and the isoparametric map affine)."""
    x = m.x.copy()
    h = np.diff(np.unique(x[:, 0]))[0] * (2 if m.eNoN == 10 else 1)
    rng = np.random.default_rng(seed)
    inner = np.all((x > 1e-12) & (x < 1 - 1e-12), axis=1)
    corners = np.unique(m.ien[:, :4])
    move = np.zeros_like(x)
    move[corners] = rng.uniform(-frac * h, frac * h, (len(corners), 3))
    move[~inner] = 0.0
    x[corners] += move[corners]
    if m.eNoN == 10:
        for k, (a, b) in enumerate(mesh._MID):
            mid = m.ien[:, 4 + k]
            x[mid] = 0.5 * (x[m.ien[:, a]] + x[m.ien[:, b]])
    out = mesh.TetMesh(x=x, ien=m.ien.copy(), node_tags={})
    assert (mesh.six_vol(out.x, out.ien) > 0).all()
    return out


def dirichlet_x0(m, nd):
    """Nodes on x = 0 with the first nd dofs fixed."""
    return np.where(m.x[:, 0] < 1e-12)[0].astype(np.int32)


def save(name, prob, rp, cp, V, R, extra=None):
    m = prob.mesh
    d = dict(x=m.x, ien=m.ien, Ag=prob.Ag, Yg=prob.Yg, Dg=prob.Dg, phys=prob.phys, dof=prob.dof, quad=prob.quad,
             eq=eq_json(prob.eq), rowPtr=rp, colPtr=cp, Val=V, R=R, prop_el=np.zeros(0),
             source="svFSI reference via oracle/ref_harness.f90")
    if extra:
        d.update(extra)
    np.savez_compressed(HERE / f"{name}.npz", **d)
    print(f"{name}: nNo {m.nNo} nEl {m.nEl} nnz {len(cp)} {(HERE / f'{name}.npz').stat().st_size // 1024} KiB")


def solve_record(prob, rp, cp, V, R, nodes, fdof, LS_type, **kw):
    """The reference's FSILS_SOLVE on (V, R) with the nodes' first fdof dofs
    masked; returns the fixture fields (prefix sol_)."""
    dv = np.ones((len(nodes), prob.dof))
    dv[:, :fdof] = 0.0
    x, st, _ = rh.solve(rp, cp, prob.dof, V, R, nodes, dv, LS_type, **kw)
    rec = dict(sol_x=x, sol_dir_nodes=nodes, sol_dir_fdof=fdof, sol_LS_type=LS_type,
               sol_settings=json.dumps({k: v for k, v in kw.items()}), sol_stats=json.dumps(st))
    print(f"   solve {LS_type}: {st}")
    return rec


def fluid_problem(n, seed, dt, umax, mu, viscType=698, **visc):
    m = jitter(mesh.cube_tet4(n), 0.2, seed)
    rng = np.random.default_rng(seed + 100)
    eq = dict(problems.gen_alpha(0.2, False), dt=dt, rho=1.06, mu=mu, bstab=0.2, fb=(0.3, -0.2, 0.1),
              viscType=viscType, **visc)
    Y = np.zeros((m.nNo, 4))
    Y[:, :3] = rng.uniform(-umax, umax, (m.nNo, 3))
    Y[:, 3] = rng.uniform(-1.0, 1.0, m.nNo)
    A = rng.normal(size=(m.nNo, 4)) * umax / dt * 1e-2
    return problems.Problem(name="fluid", phys=problems.SVFSI_PHYS_FLUID, dof=4, mesh=m, eq=eq, Ag=A, Yg=Y,
                            Dg=np.zeros_like(Y), dir_faces=[])


def fluid():
    # the benchmark's physics (rho 1.06, mu 0.04, dt 1e-3) at velocities of the inlet's scale
    p = fluid_problem(3, 1, dt=1e-3, umax=30.0, mu=0.04)
    rp, cp, V, R, _ = rh.assemble(p)
    nodes = dirichlet_x0(p.mesh, 3)
    ext = {}
    ext.update(solve_record(p, rp, cp, V, R, nodes, 3, LS_TYPE_GMRES,
                            RI=dict(relTol=1e-10, absTol=1e-30, mItr=200, sD=10)))
    save("ref_fluid_bench", p, rp, cp, V, R, ext)
    # convection-dominated: large steps and velocities, low viscosity (u' and tau_B large)
    for tag, vt, visc, mu in (("const", 698, {}, 0.004),
                              ("cy", 697, dict(mu_o=0.56, lam=3.313, visc_a=2.0, visc_n=0.3568), 0.0345),
                              ("cass", 696, dict(mu_o=0.3, lam=0.5), 0.2)):
        p = fluid_problem(3, 2, dt=0.05, umax=20.0, mu=mu, viscType=vt, **visc)
        rp, cp, V, R, _ = rh.assemble(p)
        ext = {}
        if tag == "const":
            nodes = dirichlet_x0(p.mesh, 3)
            ext.update(solve_record(p, rp, cp, V, R, nodes, 3, LS_TYPE_BICGS,
                                    RI=dict(relTol=1e-10, absTol=1e-30, mItr=500)))
            ns = solve_record(p, rp, cp, V, R, nodes, 3, LS_TYPE_NS,
                              RI=dict(relTol=1e-6, absTol=1e-30, mItr=10), GM=dict(relTol=1e-8, mItr=2, sD=20),
                              CG=dict(relTol=1e-8, mItr=500))
            ext.update({k.replace("sol_", "ns_"): v for k, v in ns.items()})
        save(f"ref_fluid_{tag}", p, rp, cp, V, R, ext)


def fluid_th_problem(n, seed, dt, umax, mu, viscType=698, **visc):
    """Taylor-Hood (P2P1) fluid on a jittered tet10 cube: random velocity,
    acceleration and corner pressure."""
    m = jitter(mesh.cube_tet10(n), 0.2, seed)
    rng = np.random.default_rng(seed + 200)
    eq = dict(problems.gen_alpha(0.2, False), dt=dt, rho=1.06, mu=mu, bstab=0.2, fb=(0.3, -0.2, 0.1),
              viscType=viscType, **visc)
    Y = np.zeros((m.nNo, 4))
    Y[:, :3] = rng.uniform(-umax, umax, (m.nNo, 3))
    corners = np.unique(m.ien[:, :4])
    Y[corners, 3] = rng.uniform(-1.0, 1.0, len(corners))
    A = rng.normal(size=(m.nNo, 4)) * umax / dt * 1e-2
    return problems.Problem(name="fluid_th", phys=problems.SVFSI_PHYS_FLUID, dof=4, mesh=m, eq=eq, Ag=A, Yg=Y,
                            Dg=np.zeros_like(Y), dir_faces=[], nFs=2)


def fluid_th():
    for tag, vt, visc, mu in (("const", 698, {}, 0.04),
                              ("cy", 697, dict(mu_o=0.56, lam=3.313, visc_a=2.0, visc_n=0.3568), 0.0345)):
        p = fluid_th_problem(2, 6, dt=0.01, umax=5.0, mu=mu, viscType=vt, **visc)
        rp, cp, V, R, _ = rh.assemble(p, nFs=2)
        save(f"ref_fluid_th10_{tag}", p, rp, cp, V, R, dict(nFs=2))


def lelas():
    m = jitter(mesh.cube_tet4(4), 0.2, 3)
    rng = np.random.default_rng(103)
    eq = dict(problems.gen_alpha(0.2, True), dt=1.0, rho=1.0, E=1e6, nu=0.3, fb=(0.0, 0.0, -9.81))
    z = np.zeros((m.nNo, 3))
    p = problems.Problem(name="lelas", phys=problems.SVFSI_PHYS_LELAS, dof=3, mesh=m, eq=eq,
                         Ag=rng.normal(size=(m.nNo, 3)), Yg=z, Dg=rng.uniform(-1e-3, 1e-3, (m.nNo, 3)),
                         dir_faces=[])
    rp, cp, V, R, _ = rh.assemble(p)
    nodes = dirichlet_x0(m, 3)
    ext = solve_record(p, rp, cp, V, R, nodes, 3, LS_TYPE_CG, RI=dict(relTol=1e-10, absTol=1e-30, mItr=2000))
    save("ref_lelas", p, rp, cp, V, R, ext)


def struct():
    for vol, tag in ((2, "st91"), (1, "quad"), (3, "m94")):
        m = jitter(mesh.cube_tet10(1), 0.15, 4)
        rng = np.random.default_rng(104 + vol)
        eq = dict(problems.gen_alpha(0.2, True), dt=1e-3, rho=1.0, C10=1e5, Kpen=1e7, volType=vol)
        p = problems.Problem(name="struct", phys=problems.SVFSI_PHYS_STRUCT, dof=3, mesh=m, eq=eq,
                             Ag=rng.normal(size=(m.nNo, 3)), Yg=rng.normal(size=(m.nNo, 3)),
                             Dg=rng.uniform(-0.05, 0.05, (m.nNo, 3)), dir_faces=[], quad=0)
        rp, cp, V, R, _ = rh.assemble(p)
        save(f"ref_struct_{tag}", p, rp, cp, V, R)


if __name__ == "__main__":
    if not rh.available():
        rh.build()
    which = sys.argv[1:] or ["fluid", "fluid_th", "lelas", "struct"]
    for name in which:
        globals()[name]()
