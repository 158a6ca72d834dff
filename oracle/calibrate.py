"""Calibration of the CPU baseline against the reference's own Fortran.

TEST INFRASTRUCTURE ONLY (it times; nothing in the product path uses it).
On the same inputs and the same host it times, for the LELAS cube of
20^3 x 6 tets with CG and the FLUID cylinder of ~200k tets with GMRES(50),
both one Newton step to a relative residual of 1e-8:

- the reference (oracle/_ref, built from the reference's sources by
  oracle/ref.mk): LHSA's pattern, CONSTRUCT_LELAS / CONSTRUCT_FLUID with
  DOASSEM, and FSILS_SOLVE, on one MPI rank with one thread (the reference's
  OpenMP pragmas are inert);
- the C oracle: the same element loop and FSILS_SOLVE (PRECONDDIAG plus
  CGRADV / GMRESV) on 1, 2, 4 and 8 threads.

The reference's multi-rank path cannot run here: DISTRIBUTE partitions the
mesh through SPLIT.c, which needs ParMETIS, so only its one-rank time is
measured.  Prints one JSON line per configuration.

    make -f oracle/ref.mk -j8 && python oracle/calibrate.py
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402
from oracle import refharness as rh  # noqa: E402
from svfsi_amd import problems  # noqa: E402


def dirichlet(prob):
    """(nodes 0-based, values (n, dof)) of the problem's Dirichlet faces:
    the product of the face masks, as PRECONDDIAG forms it."""
    m = prob.dir_mask()
    nodes = np.nonzero((m == 0.0).any(axis=1))[0].astype(np.int32)
    return nodes, m[nodes]


def calibrate(name, prob, threads=(1, 2, 4, 8)):
    rp, cp = orc.lhsa(prob.nNo, prob.mesh.ien)
    # element loop (no face terms on either side)
    faces = prob.trac_faces, prob.bflow_faces
    prob.trac_faces, prob.bflow_faces = [], []
    _, _, Vr, Rr, ref_asm = rh.assemble(prob, nrep=3)
    t0 = time.perf_counter()
    for _ in range(3):
        Vo, Ro = orc.assemble(prob, rp, cp)
    orc_asm = (time.perf_counter() - t0) / 3
    prob.trac_faces, prob.bflow_faces = faces
    asm_err = float(np.abs(Vr - Vo).max() / np.abs(Vo).max())
    # the system of the step (face terms included), solved by both
    V, R = orc.assemble(prob, rp, cp)
    nodes, dv = dirichlet(prob)
    ls = prob.ls
    RI = dict(relTol=ls["relTol"], absTol=ls["absTol"], mItr=ls["mItr"], sD=ls.get("sD", 0))
    xr, st, ref_solve = rh.solve(rp, cp, prob.dof, V, R, nodes, dv, ls["LS_type"], RI=RI)
    orc_solve = {}
    for nt in threads:
        orc.set_threads(nt)
        t0 = time.perf_counter()
        xo, _, so = orc.fsils_solve(rp, cp, prob.dof, prob.dir_mask(), ls, R, V)
        orc_solve[nt] = time.perf_counter() - t0
    orc.set_threads(1)
    rec = {
        "config": name, "nNo": prob.nNo, "nEl": prob.mesh.nEl, "nnz": int(len(cp)), "dof": prob.dof,
        "reference_1rank": {"assembly_s": round(ref_asm, 4), "solve_s": round(ref_solve, 3),
                            "itr": st["RI_itr"], "suc": st["RI_suc"]},
        "oracle": {"assembly_s_1thread": round(orc_asm, 4),
                   "solve_s": {str(k): round(v, 3) for k, v in orc_solve.items()}, "itr": so.itr, "suc": so.suc},
        "ratio_oracle_over_reference_1core": round((orc_asm + orc_solve[1]) / (ref_asm + ref_solve), 3),
        "assembly_rel_diff": asm_err,
        "solution_rel_diff": float(np.abs(xr - xo).max() / np.abs(xr).max()),
    }
    print(json.dumps(rec), flush=True)
    return rec


def main():
    if not rh.available():
        raise SystemExit("oracle/_ref is not built (make -f oracle/ref.mk)")
    calibrate("LELAS cube 20^3 x 6 tet4, CG", problems.lelas_cube(20))
    calibrate("FLUID cylinder h=0.11 tet4, GMRES(50)", problems.fluid_cylinder(h=0.11))


if __name__ == "__main__":
    main()
