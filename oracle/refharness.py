"""Driver of oracle/_ref/svfsi_ref_harness, the reference's own svFSI and
svFSILS Fortran built by oracle/ref.mk.

TEST INFRASTRUCTURE ONLY.  It is used where the reference sources are
present, to write the fixtures under tests/golden/ (make_ref_golden.py) and
to time the reference on the host; the tests themselves read the committed
fixtures, and nothing in the product path imports this module.
"""
from __future__ import annotations

import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
EXE = HERE / "_ref" / "svfsi_ref_harness"
REF_SRC = Path("/root/reference/Code/Source")

# reference codes (svFSI/CONSTS.f:78-81,103-104,188-202)
PHYS_FLUID, PHYS_STRUCT, PHYS_LELAS = 201, 202, 204
STISO_NHOOK = 604
STVOL = {1: 651, 2: 652, 3: 653}  # this package's volType -> stVol_Quad/ST91/M94
VISC_CONST, VISC_CY, VISC_CASS = 698, 697, 696
PHYS = {1: PHYS_FLUID, 2: PHYS_LELAS, 3: PHYS_STRUCT}  # this package's phys codes


def build() -> Path:
    """Compiles the reference and the harness (needs the reference sources)."""
    if not REF_SRC.is_dir():
        raise RuntimeError("the reference sources are not present")
    r = subprocess.run(["make", "-f", "oracle/ref.mk", "-j8"], cwd=ROOT, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference build failed:\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}")
    return EXE


def available() -> bool:
    return EXE.exists()


def _run(mode: str, payload: bytes) -> bytes:
    with tempfile.TemporaryDirectory() as td:
        # the run directory is temporary: NSSOLVER appends to FSILS_NS.log in it
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(payload)
        env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")
        r = subprocess.run([str(EXE), mode, fin, fout], capture_output=True, text=True, env=env, cwd=td,
                           timeout=float(os.environ.get("SVFSI_REF_TIMEOUT", "600")))
        if r.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError(f"reference harness failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        with open(fout, "rb") as f:
            return f.read()


def assemble(prob, nrep: int = 1, pS0: np.ndarray | None = None, Bf: np.ndarray | None = None,
             visc: dict | None = None, nFs: int = 1):
    """The reference's LHSA + CONSTRUCT_* for a Problem (no face terms).

    Returns (rowPtr, colPtr) 0-based, Val (nnz, dof*dof), R (nNo, dof) and the
    seconds per assembly (mean of nrep)."""
    m = prob.mesh
    eq = prob.eq
    dof = prob.dof
    visc = visc or {}
    vt = int(eq.get("viscType", VISC_CONST))
    hdr = np.zeros(16, np.int32)
    hdr[:11] = [PHYS[prob.phys], m.nNo, m.nEl, m.eNoN, nFs, vt, STVOL[int(eq.get("volType", 2))], STISO_NHOOK,
                pS0 is not None, Bf is not None, nrep]
    par = np.zeros(24)
    fb = eq.get("fb", (0.0, 0.0, 0.0))
    par[:20] = [eq["dt"], eq["am"], eq["af"], eq["beta"], eq["gam"], eq.get("rho", 0.0), eq.get("E", 0.0),
                eq.get("nu", 0.0), eq.get("mu", 0.0), eq.get("mu_o", 0.0), eq.get("lam", 0.0),
                eq.get("visc_a", 0.0), eq.get("visc_n", 0.0), eq.get("C10", 0.0), eq.get("Kpen", 0.0),
                fb[0], fb[1], fb[2], eq.get("mu_s", 0.0), eq.get("dmp", 0.0)]
    tDof = dof
    parts = [hdr.tobytes(), par.tobytes(), np.ascontiguousarray(m.x, np.float64).tobytes(),
             np.ascontiguousarray(m.ien + 1, np.int32).tobytes()]
    for a in (prob.Ag, prob.Yg, prob.Dg):
        parts.append(np.ascontiguousarray(np.asarray(a)[:, :tDof], np.float64).tobytes())
    if pS0 is not None:
        parts.append(np.ascontiguousarray(pS0, np.float64).tobytes())
    if Bf is not None:
        parts.append(np.ascontiguousarray(Bf, np.float64).tobytes())
    out = _run("asm", b"".join(parts))
    nnz = int(np.frombuffer(out, np.int32, 1)[0])
    o = 4
    rowPtr = np.frombuffer(out, np.int32, m.nNo + 1, o) - 1
    o += 4 * (m.nNo + 1)
    colPtr = np.frombuffer(out, np.int32, nnz, o) - 1
    o += 4 * nnz
    Val = np.frombuffer(out, np.float64, nnz * dof * dof, o).reshape(nnz, dof * dof)
    o += 8 * nnz * dof * dof
    R = np.frombuffer(out, np.float64, m.nNo * dof, o).reshape(m.nNo, dof)
    o += 8 * m.nNo * dof
    secs = float(np.frombuffer(out, np.float64, 1, o)[0])
    return rowPtr.astype(np.int32), colPtr.astype(np.int32), Val.copy(), R.copy(), secs


def solve(rowPtr, colPtr, dof: int, Val, R, dir_nodes, dir_val, LS_type: int, prec: int = 701,
          RI: dict | None = None, GM: dict | None = None, CG: dict | None = None, nrep: int = 1):
    """The reference's FSILS_SOLVE on one rank (0-based pattern; dir_nodes
    0-based with dir_val (n, dof) the face values: 0 masks a dof).

    Returns (x (nNo, dof), stats dict, seconds per solve)."""
    RI, GM, CG = RI or {}, GM or {}, CG or {}
    nNo = len(rowPtr) - 1
    nnz = len(colPtr)
    dn = np.asarray(dir_nodes, np.int32)
    hdr = np.zeros(16, np.int32)
    hdr[:12] = [dof, nNo, nnz, LS_type, prec, len(dn), nrep, RI.get("mItr", 0), RI.get("sD", 0),
                GM.get("mItr", 0), GM.get("sD", 0), CG.get("mItr", 0)]
    par = np.zeros(16)
    par[:6] = [RI.get("relTol", 0.0), RI.get("absTol", 0.0), GM.get("relTol", 0.0), GM.get("absTol", 0.0),
               CG.get("relTol", 0.0), CG.get("absTol", 0.0)]
    parts = [hdr.tobytes(), par.tobytes(), np.ascontiguousarray(np.asarray(rowPtr) + 1, np.int32).tobytes(),
             np.ascontiguousarray(np.asarray(colPtr) + 1, np.int32).tobytes(),
             np.ascontiguousarray(Val, np.float64).tobytes(), np.ascontiguousarray(R, np.float64).tobytes()]
    if len(dn):
        parts += [(dn + 1).astype(np.int32).tobytes(), np.ascontiguousarray(dir_val, np.float64).tobytes()]
    out = _run("solve", b"".join(parts))
    x = np.frombuffer(out, np.float64, nNo * dof).reshape(nNo, dof).copy()
    o = 8 * nNo * dof
    st = np.frombuffer(out, np.int32, 15, o)
    o += 60
    f = np.frombuffer(out, np.float64, 5, o)
    names = ("suc", "itr", "cM", "cN", "cD")
    stats = {f"RI_{k}": int(st[i]) for i, k in enumerate(names)}
    stats.update({f"GM_{k}": int(st[5 + i]) for i, k in enumerate(names)})
    stats.update({f"CG_{k}": int(st[10 + i]) for i, k in enumerate(names)})
    stats.update(iNorm=float(f[0]), fNorm=float(f[1]), dB=float(f[2]), callD=float(f[3]))
    return x, stats, float(f[4])
