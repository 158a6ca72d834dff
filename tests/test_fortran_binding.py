"""The Fortran driver's side of the boundary: svfsi_amd/fortran/fsils_gpu.f90
(iso_c_binding interfaces keeping the FSILS_* argument lists) compiled with
the image's flang and linked against libsvfsi_amd.so."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
FDIR = ROOT / "svfsi_amd" / "fortran"
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


def build_demo(tmp_path: Path) -> Path:
    if not Path(FLANG).exists():
        pytest.skip("flang not available")
    lib = ROOT / "svfsi_amd"
    subprocess.run([FLANG, "-c", str(FDIR / "fsils_gpu.f90")], cwd=tmp_path, check=True)
    exe = tmp_path / "demo"
    subprocess.run([FLANG, str(FDIR / "fsils_gpu_demo.f90"), "fsils_gpu.o", f"-L{lib}", "-lsvfsi_amd",
                    f"-Wl,-rpath,{lib}", "-o", str(exe)], cwd=tmp_path, check=True)
    return exe


def run_demo(exe: Path) -> dict:
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120).stdout
    return {l.split()[0]: l.split()[1:] for l in out.splitlines() if l.strip()}


def test_fortran_binding_host_calls(tmp_path):
    r = run_demo(build_demo(tmp_path))
    # FSILS_LS_CREATE defaults for GMRES (svFSILS/LS.f:77-82, 93-95)
    assert r["LS"] == ["0", "0.10", "4", "250"]
    # LHSA of two tets sharing a face: rows 4, 5, 5, 5, 4 -> nnz 23
    assert r["LHSA"][:2] == ["0", "23"]
    assert r["LHSA"][3:] == ["1", "5", "10", "15", "20", "24"]


@pytest.mark.gpu
def test_fortran_binding_solve_on_gpu(tmp_path):
    r = run_demo(build_demo(tmp_path))
    assert int(r["DEVICES"][0]) > 0
    stat, suc, *x = r["SOLVE"]
    assert stat == "0" and suc == "1"
    x = [float(v) for v in x]
    expect = [0.0, 0.0]
    for i in range(2, 6):
        expect += [1.0 / i, 1.0 / (2 * i)]
    assert all(abs(a - b) < 1e-5 for a, b in zip(x, expect)), x
