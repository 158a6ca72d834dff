"""The aggregation of the multigrid inner preconditioner (svfsi_amg_aggregate,
host code, no reference counterpart), on the node graphs of the LELAS tet4
cube and the tet10 cube: every node in exactly one aggregate, every
aggregate connected in the graph, no aggregate of one node unless the node
has no neighbour, sizes bounded (cap, plus the single nodes merged into the
smallest neighbour), deterministic, and the argument checks."""
import ctypes as C

import numpy as np
import pytest

from svfsi_amd import fsils, mesh
from svfsi_amd._lib import lib


def graph(m):
    rp, cp = fsils.lhsa(m.nNo, m.ien)
    return (np.ascontiguousarray(rp - 1, dtype=np.int32), np.ascontiguousarray(cp - 1, dtype=np.int32))


def aggregate(rp, cp, cap):
    n = len(rp) - 1
    agg = np.full(n, -7, dtype=np.int32)
    P = C.POINTER(C.c_int)
    na = lib().svfsi_amg_aggregate(n, rp.ctypes.data_as(P), cp.ctypes.data_as(P), cap, agg.ctypes.data_as(P))
    return na, agg


@pytest.mark.parametrize("which,cap", [("tet4", 8), ("tet4", 4), ("tet10", 8), ("tet10", 27)])
def test_aggregates_partition_the_graph_into_connected_pieces(which, cap):
    m = mesh.cube_tet4(6) if which == "tet4" else mesh.cube_tet10(3)
    rp, cp = graph(m)
    n = m.nNo
    na, agg = aggregate(rp, cp, cap)
    assert 0 < na < n
    assert agg.min() == 0 and agg.max() == na - 1
    sizes = np.bincount(agg, minlength=na)
    assert sizes.min() >= 2  # every node of these meshes has neighbours
    # a node merged alone into its smallest neighbouring aggregate adds at
    # most one member per merge; no aggregate holds more than cap + its
    # merged singles, bounded by the node degree
    assert sizes.max() <= cap + np.diff(rp).max()
    assert sizes.mean() > 1.5
    for a in range(na):
        members = np.nonzero(agg == a)[0]
        seen = {members[0]}
        stack = [members[0]]
        mset = set(members.tolist())
        while stack:
            i = stack.pop()
            for j in cp[rp[i]:rp[i + 1]]:
                if j in mset and j not in seen:
                    seen.add(j)
                    stack.append(j)
        assert len(seen) == len(members), (a, members)
    na2, agg2 = aggregate(rp, cp, cap)
    assert na2 == na and np.array_equal(agg, agg2)


def test_isolated_node_stays_alone_and_bad_input_is_refused():
    # a path 0-1-2 plus an isolated node 3 (each row lists itself)
    rp = np.array([0, 2, 5, 7, 8], dtype=np.int32)
    cp = np.array([0, 1, 0, 1, 2, 1, 2, 3], dtype=np.int32)
    na, agg = aggregate(rp, cp, 8)
    assert na == 2 and list(agg) == [0, 0, 0, 1]
    bad = cp.copy()
    bad[3] = 9
    na, _ = aggregate(rp, bad, 8)
    assert na < 0
    assert b"colPtr" in lib().svfsi_last_error()
