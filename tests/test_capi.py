"""The reference's own C API of svFSILS (FSILS_api_c.h:52-64, FSILS_c.h) on
this library: tests/capi/fsils_c_demo.c is a C driver written against those
prototypes and linked against libsvfsi_amd.so in place of svFSILS.

- Where the reference's headers are present (this container), the driver is
  compiled with the reference's FSILS_c.h included before svfsi_amd.h, so a
  C compiler checks every fsils_*_ prototype of this library against the
  reference's; a negative control shows that a changed prototype is refused.
- On the CPU: the LS defaults of FSILS_LS_CREATE, the COMMU fields of one
  rank, and the reference's stop-on-misuse behaviour.
- On the GPU: the driver solves the LELAS cube of the reference fixture
  (ref_lelas.npz, written by the reference's own Fortran) through
  fsils_solve_ and matches the reference's increment and the oracle's.
"""
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "capi" / "fsils_c_demo.c"
LIB = ROOT / "svfsi_amd"
REF_LS = Path("/root/reference/Code/Source/svFSILS")
MPI_INC = Path("/opt/conda/include")
CC = shutil.which("gcc") or shutil.which("cc")
CFLAGS = ["-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic"]


def compile_demo(out: Path, ref_header: bool) -> Path:
    if CC is None:
        pytest.skip("no C compiler")
    cmd = [CC, *CFLAGS, f"-I{ROOT / 'include'}"]
    if ref_header:
        cmd += ["-DUSE_REFERENCE_HEADER", f"-I{MPI_INC}", f"-I{REF_LS}"]
    cmd += [str(SRC), f"-L{LIB}", "-lsvfsi_amd", f"-Wl,-rpath,{LIB}", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def have_reference_headers() -> bool:
    return (REF_LS / "FSILS_c.h").is_file() and (MPI_INC / "mpi.h").is_file()


@pytest.mark.skipif(not have_reference_headers(), reason="the reference's FSILS_c.h (or mpi.h) is not present")
def test_prototypes_match_the_reference_header(tmp_path):
    compile_demo(tmp_path / "demo_ref", ref_header=True)
    # negative control: the same inclusion order refuses a prototype that
    # differs from the reference's in one parameter
    bad = tmp_path / "bad.c"
    bad.write_text('#include <mpi.h>\n#include "FSILS_c.h"\n'
                   "void fsils_solve_(void* lhs, FSILS_lsType* ls, const int* dof, double* Ri, double* Val, "
                   "int* prec, int* incL, double* res);\n")
    r = subprocess.run([CC, *CFLAGS, f"-I{MPI_INC}", f"-I{REF_LS}", "-fsyntax-only", str(bad)],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "conflicting types" in r.stderr


def parse(out: str) -> dict:
    d = {}
    for line in out.splitlines():
        w = line.split()
        if w:
            d.setdefault(w[0], []).append(w[1:])
    return d


@pytest.mark.parametrize("ref_header", [False, True])
def test_host_calls(tmp_path, ref_header):
    if ref_header and not have_reference_headers():
        pytest.skip("the reference's FSILS_c.h (or mpi.h) is not present")
    exe = compile_demo(tmp_path / "demo", ref_header)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = parse(r.stdout)
    ls = {int(v[0]): v[1:] for v in d["LS"]}
    # FSILS_LS_CREATE defaults (svFSILS/LS.f:69-95): foC, RI relTol/mItr/sD,
    # GM relTol/mItr/sD, CG relTol/mItr
    assert ls[797] == ["1", "0.1", "4", "250", "0", "0", "0", "0", "0"]
    assert ls[798] == ["1", "0.01", "1000", "0", "0", "0", "0", "0", "0"]
    assert ls[796] == ["1", "0.4", "10", "100", "0.01", "2", "100", "0.2", "500"]
    # COMMU of one rank (COMMU.f:62-81): foC masF master task tF nTasks comm
    assert d["COMMU"][0] == ["1", "1", "0", "0", "1", "1", "91"]
    assert int(d["DEVICES"][0][0]) >= 0


def test_misuse_stops_the_program(tmp_path):
    exe = compile_demo(tmp_path / "demo", ref_header=False)
    r = subprocess.run([str(exe), "fatal"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "FSILS: FATAL ERROR" in r.stderr and "LS is not created yet to be freed" in r.stderr
    assert "NOT STOPPED" not in r.stdout


def write_system(path: Path, d, nodes, LS_type, s):
    RI = s["RI"]
    nNo, dof = len(d["rowPtr"]) - 1, int(d["dof"])
    with open(path, "wb") as f:
        f.write(struct.pack("<7i", nNo, len(d["colPtr"]), dof, len(nodes), LS_type, RI["mItr"], RI.get("sD", 0)))
        f.write(struct.pack("<2d", RI["relTol"], RI.get("absTol", 1e-10)))
        f.write((d["rowPtr"] + 1).astype("<i4").tobytes())
        f.write((d["colPtr"] + 1).astype("<i4").tobytes())
        f.write((np.asarray(nodes) + 1).astype("<i4").tobytes())
        f.write(np.ascontiguousarray(d["Val"], dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(d["R"], dtype="<f8").tobytes())


@pytest.mark.gpu
def test_c_driver_solves_the_reference_lelas_cube(gpu, tmp_path, orc):
    from test_ref_golden import load, mask_of, oracle_settings, rel, say, solve_case

    p, d = load("ref_lelas")
    nodes, fdof, LS_type, s, st, xref = solve_case(d, "sol")
    assert fdof == p.dof
    exe = compile_demo(tmp_path / "demo", ref_header=False)
    write_system(tmp_path / "in.bin", d, nodes, LS_type, s)
    say("C driver: fsils_solve_ on the reference LELAS cube")
    r = subprocess.run([str(exe), "solve", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = (tmp_path / "out.bin").read_bytes()
    suc, itr = struct.unpack_from("<2i", raw, 0)
    iNorm, fNorm, dB = struct.unpack_from("<3d", raw, 8)
    x = np.frombuffer(raw, dtype="<f8", offset=32).reshape(xref.shape)
    xo, _, so = orc.fsils_solve(d["rowPtr"], d["colPtr"], p.dof, mask_of(p, nodes, fdof),
                                oracle_settings(LS_type, s), d["R"], d["Val"])
    say("itr", itr, "reference", st["RI_itr"], "oracle", so.itr, "x vs reference", rel(x, xref),
        "x vs oracle", rel(x, xo))
    assert suc == 1 == st["RI_suc"]
    assert abs(itr - st["RI_itr"]) <= 2
    assert abs(iNorm - st["iNorm"]) <= 1e-12 * st["iNorm"]
    assert rel(x, xref) < 1e-8 and rel(x, xo) < 1e-8
