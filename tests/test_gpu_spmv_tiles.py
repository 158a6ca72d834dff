"""GPU: the CSR-stream block products (row tiles of at most 256 or 512 blocks)
and the kernels used when a pattern has no valid tiling, against the
oracle's SPARMUL (SPARMUL.f:51-133) on patterns chosen for the tile edges:
rows of mixed length (tet10), rows of exactly 256 blocks (a tile is one or
two whole rows), a row of 300 blocks (no 256-block tiling, the 512 one
runs), and a row of 700 blocks (no tiling at all)."""
import sys

import numpy as np
import pytest

from svfsi_amd import fsils, mesh

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def dense_pattern(n):
    rp = 1 + np.arange(n + 1, dtype=np.int64) * n
    cp = np.tile(np.arange(1, n + 1, dtype=np.int64), n)
    return n, rp.astype(np.int32), cp.astype(np.int32)


def arrow_pattern(n):
    """row 0 couples to every node, every other row to itself and node 0"""
    rows = [np.arange(1, n + 1)] + [np.array([1, i + 1]) for i in range(1, n)]
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return n, (rp + 1).astype(np.int32), np.concatenate(rows).astype(np.int32)


def tet10_pattern():
    m = mesh.cube_tet10(4)
    rp, cp = fsils.lhsa(m.nNo, m.ien)
    return m.nNo, rp, cp


PATTERNS = {"tet10": tet10_pattern, "dense256": lambda: dense_pattern(256),
            "arrow300": lambda: arrow_pattern(300), "arrow700": lambda: arrow_pattern(700)}


@pytest.mark.parametrize("dof", [3, 4])
@pytest.mark.parametrize("pat", list(PATTERNS))
def test_spmv_on_tile_edges(gpu, orc, pat, dof):
    import torch

    nNo, rp, cp = PATTERNS[pat]()
    nnz = len(cp)
    say("spmv tiles", pat, dof, nNo, nnz, "longest row", int(np.diff(rp).max()))
    lhs = fsils.fsils_lhs_create(nNo, nnz, rp, cp)
    rng = np.random.default_rng(7 + dof)
    Val = rng.normal(size=(nnz, dof * dof))
    U = rng.normal(size=(nNo, dof))
    ref = orc.spmv(rp - 1, cp - 1, dof, Val, U)
    dV = torch.from_numpy(Val).to(gpu)
    dU = torch.from_numpy(U).to(gpu)
    out = torch.zeros((nNo, dof), dtype=torch.float64, device=gpu)
    fsils.call("svfsi_spmv_dev", lhs.handle, dof, fsils.dev_ptr(dV), fsils.dev_ptr(dU), fsils.dev_ptr(out),
               fsils.cur_stream())
    torch.cuda.synchronize()
    assert relerr(out.cpu().numpy(), ref) < 1e-13
    # the same call again gives the same bits (fixed summation order)
    out2 = torch.zeros_like(out)
    fsils.call("svfsi_spmv_dev", lhs.handle, dof, fsils.dev_ptr(dV), fsils.dev_ptr(dU), fsils.dev_ptr(out2),
               fsils.cur_stream())
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    V32, U32 = Val.astype(np.float32), U.astype(np.float32)
    ref32 = orc.spmv(rp - 1, cp - 1, dof, V32.astype(float), U32.astype(float))
    dV32 = torch.from_numpy(V32).to(gpu)
    dU32 = torch.from_numpy(U32).to(gpu)
    o32 = torch.zeros((nNo, dof), dtype=torch.float32, device=gpu)
    fsils.call("svfsi_spmv_f32_dev", lhs.handle, dof, fsils.dev_ptr(dV32), fsils.dev_ptr(dU32),
               fsils.dev_ptr(o32), fsils.cur_stream())
    torch.cuda.synchronize()
    # FP32 sums over rows of up to 700 blocks
    assert relerr(o32.cpu().numpy(), ref32) < 2e-5
    lhs.free()
