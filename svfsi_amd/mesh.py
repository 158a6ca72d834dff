"""Synthetic tetrahedral meshes for the per-Newton-step linear-system path.

Meshes are generated in memory (the reference reads them from files,
svFSI/LOADMSH.f, READMSH.f).  Element orientation follows CHECKIEN
(svFSI/READMSH.f:1133-1180): nodes are ordered so that
det[x1-x4, x2-x4, x3-x4] > 0, the sign GNN expects (svFSI/NN.f:1522-1527).
Tet10 mid-edge nodes follow GETGNN's order (svFSI/NN.f:674-715):
5:(1,2) 6:(2,3) 7:(1,3) 8:(1,4) 9:(2,4) 10:(3,4).

All connectivity returned here is 0-based.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# Kuhn split of the unit cube: one tet per axis permutation along the main
# diagonal v000 -> v111.
_PERMS = list(itertools.permutations(range(3)))
_MID = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def _orient(x: np.ndarray, ien: np.ndarray) -> np.ndarray:
    """Swap nodes 1<->2 (0-based 0<->1) where det[x1-x4,x2-x4,x3-x4] < 0."""
    p = x[ien[:, :4]]
    d = np.einsum("ij,ij->i", p[:, 0] - p[:, 3], np.cross(p[:, 1] - p[:, 3], p[:, 2] - p[:, 3]))
    neg = d < 0
    ien = ien.copy()
    ien[neg, 0], ien[neg, 1] = ien[neg, 1].copy(), ien[neg, 0].copy()
    return ien


def six_vol(x: np.ndarray, ien: np.ndarray) -> np.ndarray:
    """6 x signed volume = the GNN Jacobian of each tet4 / tet10 corner tet."""
    p = x[ien[:, :4]]
    return np.einsum("ij,ij->i", p[:, 0] - p[:, 3], np.cross(p[:, 1] - p[:, 3], p[:, 2] - p[:, 3]))


@dataclass
class TetMesh:
    x: np.ndarray            # (nNo, 3) float64
    ien: np.ndarray          # (nEl, eNoN) int32, 0-based
    node_tags: dict = field(default_factory=dict)  # name -> bool mask over nodes

    @property
    def nNo(self) -> int:
        return self.x.shape[0]

    @property
    def nEl(self) -> int:
        return self.ien.shape[0]

    @property
    def eNoN(self) -> int:
        return self.ien.shape[1]


def cube_tet4(n: int, length: float = 1.0) -> TetMesh:
    """Unit cube of n^3 hex cells, each split into 6 Kuhn tets (tet4)."""
    m = n + 1
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    nid = (ii + m * (jj + m * kk)).ravel()
    x = np.zeros((m ** 3, 3))
    x[nid, 0] = ii.ravel() * length / n
    x[nid, 1] = jj.ravel() * length / n
    x[nid, 2] = kk.ravel() * length / n
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")
    tets = []
    for p in _PERMS:
        v = [np.zeros(3, int)]
        for ax in p[:2]:
            nxt = v[-1].copy()
            nxt[ax] += 1
            v.append(nxt)
        v.append(np.ones(3, int))
        tets.append(np.stack([(ci + a[0]) + m * ((cj + a[1]) + m * (ck + a[2])) for a in v], axis=1))
    ien = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    ien = _orient(x, ien)
    return TetMesh(x=x, ien=ien.astype(np.int32))


def cube_tet10(n: int, length: float = 1.0) -> TetMesh:
    """Kuhn cube of n^3 cells x 6 quadratic tets; nodes on the (2n+1)^3 lattice."""
    m = 2 * n + 1
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    nid = (ii + m * (jj + m * kk)).ravel()
    x = np.zeros((m ** 3, 3))
    x[nid, 0] = ii.ravel() * length / (2 * n)
    x[nid, 1] = jj.ravel() * length / (2 * n)
    x[nid, 2] = kk.ravel() * length / (2 * n)
    lin = cube_tet4(n, length)
    # corner ids on the doubled lattice
    mm = n + 1
    c = lin.ien.astype(np.int64)
    ci, cj, ck = c % mm, (c // mm) % mm, c // (mm * mm)
    corner = (2 * ci) + m * ((2 * cj) + m * (2 * ck))
    mids = []
    for a, b in _MID:
        mi = ci[:, a] + ci[:, b]
        mj = cj[:, a] + cj[:, b]
        mk = ck[:, a] + ck[:, b]
        mids.append(mi + m * (mj + m * mk))
    ien = np.concatenate([corner, np.stack(mids, axis=1)], axis=1).astype(np.int32)
    return TetMesh(x=x, ien=ien)


def boundary_faces(ien: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Corner triangles of tets that belong to exactly one element.

    Returns (faces (nF,3) 0-based, in the element's node order, and the
    adjacent element id of each face)."""
    c = ien[:, :4].astype(np.int64)
    loc = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])
    f = c[:, loc].reshape(-1, 3)
    el = np.repeat(np.arange(ien.shape[0]), 4)
    s = np.sort(f, axis=1)
    big = int(s.max()) + 1
    key = (s[:, 0] * big + s[:, 1]) * big + s[:, 2]
    uk, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    once = cnt[inv] == 1
    return f[once].astype(np.int32), el[once].astype(np.int32)


def cylinder_tet4(h: float, radius: float = 1.0, length: float = 10.0, seed: int = 0,
                  jitter: float = 0.2) -> TetMesh:
    """Straight cylinder (axis z) from a Delaunay tetrahedralisation of seeded,
    jittered lattice points: interior lattice, wall rings at r = radius, and
    jittered planar points on the two caps.  Node tags: 'inlet' (z = 0),
    'outlet' (z = length), 'wall' (r = radius)."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    nz = max(2, int(round(length / h)))
    hz = length / nz
    nxy = int(np.ceil(radius / h))
    g = np.arange(-nxy, nxy + 1) * h
    # interior
    X, Y, Z = np.meshgrid(g, g, np.arange(1, nz) * hz, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    P = P + rng.uniform(-jitter, jitter, P.shape) * h
    r = np.hypot(P[:, 0], P[:, 1])
    P = P[(r < radius - 0.5 * h) & (P[:, 2] > 0.4 * hz) & (P[:, 2] < length - 0.4 * hz)]
    # wall rings
    nth = max(8, int(round(2 * np.pi * radius / h)))
    rings = []
    for k in range(nz + 1):
        th = 2 * np.pi * (np.arange(nth) + rng.uniform(-jitter, jitter, nth)) / nth + rng.uniform(0, 2 * np.pi)
        rings.append(np.stack([radius * np.cos(th), radius * np.sin(th), np.full(nth, k * hz)], 1))
    W = np.concatenate(rings)
    W[:, 2] = np.where(np.isclose(W[:, 2], length), length, W[:, 2])
    # caps
    X2, Y2 = np.meshgrid(g, g, indexing="ij")
    Q = np.stack([X2.ravel(), Y2.ravel()], 1)
    caps = []
    for zc in (0.0, length):
        q = Q + rng.uniform(-jitter, jitter, Q.shape) * h
        q = q[np.hypot(q[:, 0], q[:, 1]) < radius - 0.5 * h]
        caps.append(np.concatenate([q, np.full((len(q), 1), zc)], 1))
    pts = np.concatenate([P, W] + caps)
    tri = Delaunay(pts)
    ien = tri.simplices.astype(np.int64)
    vol6 = six_vol(pts, ien)
    keep = np.abs(vol6) > 1e-9 * h ** 3
    ien = ien[keep]
    # drop unused points, renumber by (z, y, x) slab order for locality
    used = np.unique(ien)
    pts = pts[used]
    remap = np.full(int(used.max()) + 1, -1, np.int64)
    remap[used] = np.arange(len(used))
    ien = remap[ien]
    order = np.lexsort((pts[:, 0], pts[:, 1], np.floor(pts[:, 2] / hz + 0.5)))
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    pts = pts[order]
    ien = inv[ien]
    ien = ien[np.argsort(ien.min(axis=1), kind="stable")]
    ien = _orient(pts, ien.astype(np.int32))
    rr = np.hypot(pts[:, 0], pts[:, 1])
    tags = {
        "inlet": pts[:, 2] == 0.0,
        "outlet": pts[:, 2] == length,
        "wall": np.abs(rr - radius) < 1e-12,
    }
    return TetMesh(x=pts, ien=ien.astype(np.int32), node_tags=tags)
