"""Sparsity pattern (LHSA, svFSI/LHSA.f:38-262): the oracle against brute
force, and the product library's svfsi_lhsa against the oracle, bit for bit
(host routine: no GPU needed)."""
import numpy as np
import pytest

from svfsi_amd import fsils, mesh


def brute(nNo, ien):
    rows = [set() for _ in range(nNo)]
    for el in ien:
        for a in el:
            rows[a].update(int(b) for b in el)
    rp = [0]
    cp = []
    for r in rows:
        cp += sorted(r)
        rp.append(len(cp))
    return np.array(rp), np.array(cp)


@pytest.mark.parametrize("mk", [lambda: mesh.cube_tet4(3), lambda: mesh.cube_tet10(2),
                                lambda: mesh.cylinder_tet4(0.6, length=2.0)])
def test_oracle_lhsa_bruteforce(orc, mk):
    m = mk()
    rp, cp = orc.lhsa(m.nNo, m.ien)
    brp, bcp = brute(m.nNo, m.ien)
    assert np.array_equal(rp, brp)
    assert np.array_equal(cp, bcp)
    # diagonal present, sorted unique
    for r in range(m.nNo):
        row = cp[rp[r]:rp[r + 1]]
        assert r in row
        assert np.all(np.diff(row) > 0)


def test_oracle_lhsa_detects_isolated_node(orc):
    m = mesh.cube_tet4(1)
    with pytest.raises(ValueError):
        orc.lhsa(m.nNo + 1, m.ien)


@pytest.mark.parametrize("mk", [lambda: mesh.cube_tet4(20), lambda: mesh.cube_tet10(4),
                                lambda: mesh.cylinder_tet4(0.3)])
def test_library_lhsa_bitexact(orc, mk):
    m = mk()
    rp, cp = orc.lhsa(m.nNo, m.ien)
    lrp, lcp = fsils.lhsa(m.nNo, m.ien)
    assert np.array_equal(lrp - 1, rp)
    assert np.array_equal(lcp - 1, cp)


def test_config_sizes():
    # LELAS cube 20^3 x 6: 9,261 nodes, 48,000 tets, 128,581 blocks
    m = mesh.cube_tet4(20)
    rp, cp = fsils.lhsa(m.nNo, m.ien)
    assert (m.nNo, m.nEl, len(cp)) == (9261, 48000, 128581)
    # 55^3 x 6 tet4 cube: 175,616 nodes, 998,250 tets, 2,559,646 blocks
    m = mesh.cube_tet4(55)
    rp, cp = fsils.lhsa(m.nNo, m.ien)
    assert (m.nNo, m.nEl, len(cp)) == (175616, 998250, 2559646)
