"""Regression fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle on tiny seeded inputs): the oracle must keep reproducing them, the
library's LHSA must reproduce their patterns, and the HIP path must match
them (GPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

from svfsi_amd import fsils, mesh, problems

GOLD = Path(__file__).resolve().parent / "golden"
NAMES = ["lelas_cube2", "fluid_cube1", "struct_cube1_q0", "struct_cube1_q1"]


def load(name):
    d = np.load(GOLD / f"{name}.npz", allow_pickle=False)
    m = mesh.TetMesh(x=d["x"], ien=d["ien"], node_tags={})
    prop = d["prop_el"] if d["prop_el"].size else None
    p = problems.Problem(name=name, phys=int(d["phys"]), dof=int(d["dof"]), mesh=m, eq=json.loads(str(d["eq"])),
                         Ag=d["Ag"], Yg=d["Yg"], Dg=d["Dg"], dir_faces=[], prop_el=prop, quad=int(d["quad"]))
    if "trac_fien" in d:
        p.trac_faces = [(d["trac_fien"], d["trac_fel"], d["trac_h"])]
    return p, d


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(orc, name):
    p, d = load(name)
    rp, cp = orc.lhsa(p.nNo, p.mesh.ien)
    assert np.array_equal(rp, d["rowPtr"]) and np.array_equal(cp, d["colPtr"])
    V, R = orc.assemble(p, rp, cp)
    assert rel(V, d["Val"]) < 1e-13 and rel(R, d["R"]) < 1e-13
    if "g" in d and p.phys == problems.SVFSI_PHYS_FLUID:
        assert rel(orc.fluid_dmu_dot(p, d["lam"], d["du"]), d["g"]) < 1e-13
    if "g" in d and p.phys == problems.SVFSI_PHYS_LELAS:
        assert rel(orc.lelas_dkde_dot(p, d["lam"], d["sol"]), d["g"]) < 1e-13


@pytest.mark.parametrize("name", NAMES)
def test_library_lhsa_reproduces_fixture(name):
    p, d = load(name)
    rp, cp = fsils.lhsa(p.nNo, p.mesh.ien)
    assert np.array_equal(rp - 1, d["rowPtr"]) and np.array_equal(cp - 1, d["colPtr"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_assembly_matches_fixture(gpu, name):
    import torch

    from svfsi_amd.step import NewtonStep

    p, d = load(name)
    st = NewtonStep(p)
    st.assemble()
    torch.cuda.synchronize()
    assert rel(st.Val.cpu().numpy(), d["Val"]) < 1e-12
    assert rel(st.R.cpu().numpy(), d["R"]) < 1e-12
    if p.phys == problems.SVFSI_PHYS_FLUID:
        g = torch.zeros(p.mesh.nEl, dtype=torch.float64, device=gpu)
        lam, du = (torch.from_numpy(d[k]).to(gpu) for k in ("lam", "du"))
        st.dmesh.fluid_dmu_dot(p.eq, st.Ag, st.Yg, st.prop, lam, du, g)
        torch.cuda.synchronize()
        assert rel(g.cpu().numpy(), d["g"]) < 1e-12
    if p.phys == problems.SVFSI_PHYS_LELAS:
        g = torch.zeros(p.mesh.nEl, dtype=torch.float64, device=gpu)
        lam, x = (torch.from_numpy(d[k]).to(gpu) for k in ("lam", "sol"))
        st.dmesh.lelas_dkde_dot(p.eq, lam, x, g)
        torch.cuda.synchronize()
        assert rel(g.cpu().numpy(), d["g"]) < 1e-12
