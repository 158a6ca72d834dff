"""Mesh files on the input side of the path, through the C ABI:
READCCNE's coordinate/connectivity files and face (ebc) files
(svFSI/LOADMSH.f:125-239) and CHECKIEN's orientation fix
(svFSI/READMSH.f:1010-1185)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call
from .mesh import TetMesh


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_ccne(conn_path: str, coord_path: str, check: bool = True) -> tuple[np.ndarray, np.ndarray, int]:
    """(ien (nEl, eNoN) 0-based, x (nNo, nsd), elements reordered by
    CHECKIEN) from a connectivity and a coordinates file."""
    n = [C.c_int() for _ in range(4)]
    call("svfsi_read_ccne_sizes", conn_path.encode(), coord_path.encode(), *[C.byref(v) for v in n])
    nEl, eNoN, nNo, nsd = (v.value for v in n)
    ien = np.zeros((nEl, eNoN), dtype=np.int32)
    x = np.zeros((nNo, nsd))
    call("svfsi_read_ccne", conn_path.encode(), coord_path.encode(), nEl, eNoN, nNo, nsd, _ip(ien), _dp(x))
    swapped = 0
    if check:  # ien is 1-based, as read
        if nsd != 3:
            raise ValueError("CHECKIEN is implemented for 3-D tet meshes")
        ien, swapped = checkien(ien, x)
    return ien - 1, x, swapped


def checkien(ien1: np.ndarray, x: np.ndarray) -> tuple[np.ndarray, int]:
    """CHECKIEN on a 1-based (nEl, eNoN) tet4/tet10 connectivity: returns the
    reordered copy and the number of elements it reordered."""
    ien = np.ascontiguousarray(ien1, dtype=np.int32).copy()
    xx = np.ascontiguousarray(x, dtype=np.float64)
    k = C.c_int()
    call("svfsi_checkien", ien.shape[1], ien.shape[0], xx.shape[0], _ip(ien), _dp(xx), C.byref(k))
    return ien, k.value


def read_ebc(path: str) -> tuple[np.ndarray, np.ndarray]:
    """(element ids gE (1-based, as read), face nodes (nFe, feNoN) 0-based)."""
    n = [C.c_int(), C.c_int()]
    call("svfsi_read_ebc_sizes", path.encode(), C.byref(n[0]), C.byref(n[1]))
    nFe, feNoN = n[0].value, n[1].value
    gE = np.zeros(nFe, dtype=np.int32)
    fien = np.zeros((nFe, feNoN), dtype=np.int32)
    call("svfsi_read_ebc", path.encode(), nFe, feNoN, _ip(gE), _ip(fien))
    return gE, fien - 1


def load_mesh(conn_path: str, coord_path: str) -> TetMesh:
    """A TetMesh from READCCNE files, CHECKIEN applied."""
    ien, x, _ = read_ccne(conn_path, coord_path)
    return TetMesh(x=x, ien=ien)


def read_vtk(path: str):
    """(x (nPoints, 3), cells (nCells, nodesPerCell) 0-based, GlobalNodeID or
    None, GlobalElementID or None) from a .vtu or .vtp file."""
    n = [C.c_int() for _ in range(5)]
    call("svfsi_read_vtk_sizes", path.encode(), *[C.byref(v) for v in n])
    nP, nC, npc, hasN, hasE = (v.value for v in n)
    x = np.zeros((nP, 3))
    conn = np.zeros((nC, npc), dtype=np.int32)
    gn = np.zeros(nP, dtype=np.int32) if hasN else None
    ge = np.zeros(nC, dtype=np.int32) if hasE else None
    call("svfsi_read_vtk", path.encode(), nP, nC, npc, _dp(x), _ip(conn), None if gn is None else _ip(gn),
         None if ge is None else _ip(ge))
    return x, conn, gn, ge
