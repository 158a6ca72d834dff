"""The C-ABI library loads and exports every function include/svfsi_amd.h
declares; without a GPU the compute entry points fail loudly (no CPU
fallback)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from svfsi_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "svfsi_amd.h"


def declared():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "fsils_gpu_solve" in names and "svfsi_assemble_dev" in names
    assert len(names) >= 25


def test_library_exports_all_declared():
    L = _lib.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers them
    assert set(declared()) <= set(_lib.SIGNATURES) | set(_lib.VOID_SIGNATURES) | {
        "svfsi_last_error", "svfsi_noalloc_scratch_bytes"}


def test_ls_create_defaults_match_reference():
    from svfsi_amd import fsils

    ls = fsils.fsils_ls_create(fsils.LS_TYPE_GMRES)
    assert (ls.RI.relTol, ls.RI.mItr, ls.RI.sD, ls.RI.absTol) == (0.1, 4, 250, 1e-10)
    ls = fsils.fsils_ls_create(fsils.LS_TYPE_CG, relTol=1e-8, maxItr=7)
    assert (ls.RI.relTol, ls.RI.mItr, ls.RI.absTol) == (1e-8, 7, 1e-10)
    with pytest.raises(_lib.SvfsiError):
        fsils.fsils_ls_create(123)


def test_no_cpu_fallback_without_device():
    from svfsi_amd import fsils

    if fsils.device_count() > 0:
        pytest.skip("a GPU is present")
    rp = np.array([1, 2], np.int32)
    cp = np.array([1], np.int32)
    with pytest.raises(_lib.SvfsiError, match="no HIP device"):
        fsils.fsils_lhs_create(1, 1, rp, cp)
