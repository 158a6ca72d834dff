"""Python mirror of the svFSILS interface, over the libsvfsi_amd C ABI.

Names, argument meaning and error behaviour follow svFSILS/FSILS_API.h:51-135
(FSILS_LHS_CREATE, FSILS_BC_CREATE, FSILS_LS_CREATE, FSILS_SOLVE).  Misuse
raises SvfsiError where the reference STOPs with "FSILS: FATAL ERROR";
non-convergence is reported through ls.RI.suc as in the reference.

Arrays use the reference memory layout: Ri(dof,nNo) is a C array of shape
(nNo, dof), Val(dof*dof,nnz) a C array of shape (nnz, dof*dof); index
arrays crossing the ABI (rowPtr, colPtr, gNodes) are 1-based int32.
Host arrays are numpy; device arrays are torch CUDA tensors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (BC_TYPE_Dir, BC_TYPE_Neu, FSILS_lsType, LS_TYPE_BICGS, LS_TYPE_CG, LS_TYPE_GMRES, LS_TYPE_NS,
                   PRECOND_FSILS,
                   PRECOND_RCS, SvfsiError, call, lib)

__all__ = [
    "Lhs", "fsils_lhs_create", "fsils_bc_create", "fsils_ls_create", "fsils_solve", "lhsa",
    "BC_TYPE_Dir", "BC_TYPE_Neu", "LS_TYPE_BICGS", "LS_TYPE_CG", "LS_TYPE_GMRES", "LS_TYPE_NS", "PRECOND_FSILS",
    "PRECOND_RCS",
]


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def dev_ptr(t) -> C.c_void_p:
    """Raw device pointer of a contiguous torch CUDA tensor."""
    if not t.is_cuda:
        raise SvfsiError("expected a CUDA tensor")
    if not t.is_contiguous():
        raise SvfsiError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def cur_stream() -> C.c_void_p:
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def lhsa(nNo: int, ien: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Block sparsity pattern of a mesh (LHSA, svFSI/LHSA.f:38-262).

    ien: (nEl, eNoN) 0-based.  Returns 1-based rowPtr (nNo+1) and colPtr."""
    ien1 = _i32(ien) + 1
    nEl, eNoN = ien1.shape
    rowPtr = np.zeros(nNo + 1, np.int32)
    nnz = C.c_int(0)
    call("svfsi_lhsa", nNo, nEl, eNoN, _iptr(ien1), C.byref(nnz), _iptr(rowPtr), None)
    colPtr = np.zeros(nnz.value, np.int32)
    call("svfsi_lhsa", nNo, nEl, eNoN, _iptr(ien1), C.byref(nnz), _iptr(rowPtr), _iptr(colPtr))
    return rowPtr, colPtr


class Lhs:
    """FSILS_lhsType handle: the block-CSR pattern resident on the device."""

    def __init__(self, handle: C.c_void_p, nNo: int, nnz: int, nFaces: int,
                 rowPtr: np.ndarray, colPtr: np.ndarray):
        self.handle = handle
        self.nNo = nNo
        self.nnz = nnz
        self.nFaces = nFaces
        self.rowPtr = rowPtr
        self.colPtr = colPtr

    def free(self) -> None:
        if self.handle:
            call("fsils_gpu_lhs_free", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fsils_lhs_create(nNo: int, nnz: int, rowPtr, colPtr, nFaces: int = 0,
                     gnNo: int | None = None, gNodes=None) -> Lhs:
    """FSILS_LHS_CREATE (svFSILS/LHS.f:51-111): 1-based rowPtr/colPtr."""
    rowPtr = _i32(rowPtr)
    colPtr = _i32(colPtr)
    if gNodes is None:
        gNodes = np.arange(1, nNo + 1, dtype=np.int32)
    gNodes = _i32(gNodes)
    h = C.c_void_p()
    call("fsils_gpu_lhs_create", C.byref(h), gnNo if gnNo is not None else nNo, nNo, nnz,
         _iptr(gNodes), _iptr(rowPtr), _iptr(colPtr), nFaces)
    return Lhs(h, nNo, nnz, nFaces, rowPtr, colPtr)


def fsils_bc_create(lhs: Lhs, faIn: int, nNo: int, dof: int, BC_type: int, gNodes, Val=None) -> None:
    """FSILS_BC_CREATE (svFSILS/BC.f:50-121): faIn and gNodes 1-based;
    Val(dof,nNo) defaults to 0 (homogeneous Dirichlet mask)."""
    g = _i32(gNodes)
    v = None if Val is None else _f64(Val)
    call("fsils_gpu_bc_create", lhs.handle, faIn, nNo, dof, BC_type, _iptr(g),
         None if v is None else _dptr(v))


def fsils_ls_create(LS_type: int, relTol=None, absTol=None, maxItr=None, dimKry=None,
                    relTolIn=None, absTolIn=None, maxItrIn=None) -> FSILS_lsType:
    """FSILS_LS_CREATE (svFSILS/LS.f:50-119) with the reference defaults."""
    ls = FSILS_lsType()

    def dp(v, n=1):
        if v is None:
            return None
        arr = (C.c_double * n)(*(np.atleast_1d(v).tolist()))
        return C.cast(arr, C.POINTER(C.c_double)), arr

    def ip(v, n=1):
        if v is None:
            return None
        arr = (C.c_int * n)(*(np.atleast_1d(v).astype(int).tolist()))
        return C.cast(arr, C.POINTER(C.c_int)), arr

    keep = []
    args = []
    for v, f, n in ((relTol, dp, 1), (absTol, dp, 1), (maxItr, ip, 1), (dimKry, ip, 1),
                    (relTolIn, dp, 2), (absTolIn, dp, 2), (maxItrIn, ip, 2)):
        r = f(v, n)
        if r is None:
            args.append(None)
        else:
            args.append(r[0])
            keep.append(r[1])
    call("fsils_gpu_ls_create", C.byref(ls), LS_type, *args)
    return ls


def fsils_solve(lhs: Lhs, ls: FSILS_lsType, dof: int, Ri, Val, prec: int = PRECOND_FSILS,
                incL=None, res=None, mixed: bool = False, stream=None) -> None:
    """FSILS_SOLVE (svFSILS/SOLVE.f:51-143).

    Ri (nNo, dof): RHS on entry, solution on exit.  Val (nnz, dof*dof): left
    scaled in place (W K W), as the reference does.  numpy arrays use the
    host entry point; torch CUDA tensors the device-resident one.
    mixed=True selects the mixed-precision solve (device arrays only)."""
    inc = None if incL is None else _i32(incL)
    rs = None if res is None else _f64(res)
    incp = None if inc is None else _iptr(inc)
    resp = None if rs is None else _dptr(rs)
    if isinstance(Ri, np.ndarray):
        if mixed:
            raise SvfsiError("the mixed-precision solve takes device arrays")
        if Ri.dtype != np.float64 or not Ri.flags.c_contiguous or Val.dtype != np.float64 \
                or not Val.flags.c_contiguous:
            raise SvfsiError("Ri and Val must be C-contiguous float64 arrays")
        call("fsils_gpu_solve", lhs.handle, C.byref(ls), dof, _dptr(Ri), _dptr(Val), prec, incp, resp)
        return
    import torch

    if Ri.dtype != torch.float64 or Val.dtype != torch.float64:
        raise SvfsiError("Ri and Val must be float64 tensors")
    s = cur_stream() if stream is None else stream
    name = "fsils_gpu_solve_mixed_dev" if mixed else "fsils_gpu_solve_dev"
    call(name, lhs.handle, C.byref(ls), dof, dev_ptr(Ri), dev_ptr(Val), prec, incp, resp, s)


def fsils_solve_t(lhs: Lhs, ls: FSILS_lsType, dof: int, Val, rhs, lam, incL=None, stream=None,
                  mixed: bool = False) -> None:
    """Transposed solve (W K^T W) y = W rhs, lam = W y, with the forward
    solve's scaling and Dirichlet mask.  Val is the unscaled matrix and is
    not modified.  Device tensors only; mixed=True uses the mixed-precision
    solver."""
    inc = None if incL is None else _i32(incL)
    s = cur_stream() if stream is None else stream
    name = "fsils_gpu_solve_t_mixed_dev" if mixed else "fsils_gpu_solve_t_dev"
    call(name, lhs.handle, C.byref(ls), dof, dev_ptr(Val), dev_ptr(rhs), dev_ptr(lam),
         None if inc is None else _iptr(inc), s)


def device_count() -> int:
    return lib().svfsi_device_count()
