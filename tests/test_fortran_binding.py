"""The Fortran driver's side of the boundary: svfsi_amd/fortran/fsils_gpu.f90
(iso_c_binding interfaces: the fsils_gpu_* entry points, the reference's C
API fsils_*_, the mesh files, the assembly entry points and device memory)
compiled with the image's flang and linked against libsvfsi_amd.so.

svfsi_amd/fortran/fsils_gpu_demo.f90 is the hook sequence a maintainer adds
to svFSI (the one printed in the integration guide).  On the CPU it compiles
and links and makes the host calls; on the GPU it reads a LELAS cube from
coordinate/connectivity and face files, assembles it with the traction face
and solves it, and the result matches the oracle's assembly and solve of the
same problem."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from svfsi_amd import mesh, problems

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


def run_demo(exe: Path, *args) -> dict:
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        w = line.split()
        if w:
            out.setdefault(w[0], []).append(w[1:])
    return out


def test_fortran_binding_host_calls(tmp_path):
    r = run_demo(build_demo(tmp_path))
    # FSILS_LS_CREATE defaults for GMRES (svFSILS/LS.f:77-82, 93-95)
    assert r["LS"][0] == ["0", "0.10", "4", "250"]
    # LHSA of two tets sharing a face: rows 4, 5, 5, 5, 4 -> nnz 23
    assert r["LHSA"][0][:2] == ["0", "23"]
    assert r["LHSA"][0][3:] == ["1", "5", "10", "15", "20", "24"]


@pytest.mark.gpu
def test_fortran_binding_solve_on_gpu(gpu, tmp_path):
    r = run_demo(build_demo(tmp_path))
    assert int(r["DEVICES"][0][0]) > 0
    stat, suc, *x = r["SOLVE"][0]
    assert stat == "0" and suc == "1"
    x = [float(v) for v in x]
    expect = [0.0, 0.0]
    for i in range(2, 6):
        expect += [1.0 / i, 1.0 / (2 * i)]
    assert all(abs(a - b) < 1e-5 for a, b in zip(x, expect)), x


def write_lelas_files(d: Path, p: problems.Problem):
    m = p.mesh
    np.savetxt(d / "conn.txt", np.hstack([np.arange(1, m.nEl + 1)[:, None], m.ien + 1]), fmt="%d")
    np.savetxt(d / "coord.txt", np.hstack([np.arange(1, m.nNo + 1)[:, None], m.x]), fmt=["%d"] + ["%.17g"] * 3)
    for name, pred in (("x0.ebc", m.x[:, 0] == 0.0), ("x1.ebc", m.x[:, 0] == 1.0)):
        f, e = problems._faces_where(m, pred)
        np.savetxt(d / name, np.hstack([e[:, None] + 1, np.arange(1, len(e) + 1)[:, None], f + 1]), fmt="%d")


@pytest.mark.gpu
def test_fortran_driver_assembles_and_solves(gpu, tmp_path, orc):
    p = problems.lelas_cube(4)
    write_lelas_files(tmp_path, p)
    exe = build_demo(tmp_path)
    print("Fortran hooks: LELAS cube 4^3 x 6", flush=True)
    r = run_demo(exe, tmp_path)
    nNo, nEl, eNoN, nRe = map(int, r["MESH"][0])
    assert (nNo, nEl, eNoN, nRe) == (p.nNo, p.mesh.nEl, 4, 0)
    suc, itr = map(int, r["STAT"][0][:2])
    x = np.array([[float(v) for v in row[1:]] for row in r["SOL"]])
    rp, cp = orc.lhsa(p.nNo, p.mesh.ien)
    Vo, Ro = orc.assemble(p, rp, cp)
    xo, _, so = orc.fsils_solve(rp, cp, p.dof, p.dir_mask(), dict(p.ls, relTol=1e-10, mItr=2000), Ro, Vo)
    err = np.abs(x - xo).max() / np.abs(xo).max()
    print("itr", itr, "oracle", so.itr, "rel err", err, flush=True)
    assert suc == 1 and so.suc == 1 and abs(itr - so.itr) <= 2
    assert err < 1e-8
