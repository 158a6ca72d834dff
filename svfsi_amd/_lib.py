"""ctypes binding of libsvfsi_amd.so (declarations in include/svfsi_amd.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing the import of the binding raises, so nothing
can silently run on the CPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsvfsi_amd.so"

LS_TYPE_CG = 798
LS_TYPE_GMRES = 797
LS_TYPE_NS = 796
LS_TYPE_BICGS = 795
PRECOND_FSILS = 701
PRECOND_RCS = 709
BC_TYPE_Dir = 0
BC_TYPE_Neu = 1

SVFSI_PHYS_FLUID = 1
SVFSI_PHYS_LELAS = 2
SVFSI_PHYS_STRUCT = 3

# FLUID on tet10: the Taylor-Hood (P2P1, nFs = 2) basis, passed as quad
SVFSI_FLUID_P2P1 = 0x400

# STRUCT element kernel, ORed into the assembly's quad argument
SVFSI_STRUCT_AUTO = 0x000
SVFSI_STRUCT_VALU = 0x100
SVFSI_STRUCT_MFMA64 = 0x200
SVFSI_STRUCT_MFMA32 = 0x300

SVFSI_OK = 0
SVFSI_ERR_ARG = -1
SVFSI_ERR_HIP = -2
SVFSI_ERR_STATE = -3
SVFSI_ERR_NOTSUP = -4
SVFSI_ERR_NODEVICE = -5


class FSILS_subLsType(C.Structure):
    """FSILS_subLsType (svFSILS/FSILS_struct_c.h:115-130)."""

    _fields_ = [
        ("suc", C.c_int), ("mItr", C.c_int), ("sD", C.c_int), ("itr", C.c_int),
        ("cM", C.c_int), ("cN", C.c_int), ("cD", C.c_int), ("reserve", C.c_int),
        ("absTol", C.c_double), ("relTol", C.c_double), ("iNorm", C.c_double),
        ("fNorm", C.c_double), ("dB", C.c_double), ("callD", C.c_double),
    ]


class FSILS_lsType(C.Structure):
    """FSILS_lsType (svFSILS/FSILS_struct_c.h:132-140)."""

    _fields_ = [
        ("foC", C.c_int), ("LS_type", C.c_int), ("Resm", C.c_int), ("Resc", C.c_int),
        ("GM", FSILS_subLsType), ("CG", FSILS_subLsType), ("RI", FSILS_subLsType),
    ]


class FSILS_commuType(C.Structure):
    """FSILS_commuType (svFSILS/FSILS_struct_c.h:60-68); comm is an MPI_Aint."""

    _fields_ = [
        ("foC", C.c_int), ("masF", C.c_int), ("master", C.c_int), ("task", C.c_int),
        ("tF", C.c_int), ("nTasks", C.c_int), ("comm", C.c_ssize_t),
    ]


class EqParams(C.Structure):
    """svfsi_eq_params: eq(cEq) and domain properties of COMMOD."""

    _fields_ = [
        ("am", C.c_double), ("af", C.c_double), ("beta", C.c_double), ("gam", C.c_double),
        ("dt", C.c_double), ("rho", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("fb", C.c_double * 3), ("mu", C.c_double), ("bstab", C.c_double),
        ("C10", C.c_double), ("Kpen", C.c_double), ("volType", C.c_int), ("pad", C.c_int),
        ("viscType", C.c_int), ("pad2", C.c_int), ("mu_o", C.c_double), ("lam", C.c_double),
        ("visc_a", C.c_double), ("visc_n", C.c_double),
    ]


SVFSI_VISC_CONST = 698
SVFSI_VISC_CY = 697
SVFSI_VISC_CASS = 696

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> argtypes ; every entry point returns int (0 = success)
SIGNATURES = {
    "svfsi_device_count": [],
    "fsils_gpu_ls_create": [C.POINTER(FSILS_lsType), _I, _DP, _DP, _IP, _IP, _DP, _DP, _IP],
    "fsils_gpu_lhs_create": [C.POINTER(_P), _I, _I, _I, _IP, _IP, _IP, _I],
    "fsils_gpu_lhs_free": [_P],
    "fsils_gpu_bc_create": [_P, _I, _I, _I, _I, _IP, _DP],
    "fsils_gpu_bc_free": [_P, _I],
    "fsils_gpu_solve": [_P, C.POINTER(FSILS_lsType), _I, _DP, _DP, _I, _IP, _DP],
    "fsils_gpu_solve_dev": [_P, C.POINTER(FSILS_lsType), _I, _P, _P, _I, _IP, _DP, _P],
    "fsils_gpu_solve_mixed_dev": [_P, C.POINTER(FSILS_lsType), _I, _P, _P, _I, _IP, _DP, _P],
    "fsils_gpu_solve_t_dev": [_P, C.POINTER(FSILS_lsType), _I, _P, _P, _P, _IP, _P],
    "fsils_gpu_solve_t_mixed_dev": [_P, C.POINTER(FSILS_lsType), _I, _P, _P, _P, _IP, _P],
    "svfsi_spmv_dev": [_P, _I, _P, _P, _P, _P],
    "svfsi_spmv_f32_dev": [_P, _I, _P, _P, _P, _P],
    "svfsi_spmv_t_dev": [_P, _I, _P, _P, _P, _P],
    "svfsi_resid_dd_dev": [_P, _I, _P, _P, _P, _P, _I, _P],
    "svfsi_amg_aggregate": [_I, _P, _P, _I, _P],
    "svfsi_amg_levels": [_P, _P, _I],
    "svfsi_dot_dev": [_I, _P, _P, _DP, _P],
    "svfsi_precond_diag_dev": [_P, _I, _P, _P, _P, _IP, _P],
    "svfsi_precond_rcs_dev": [_P, _I, _P, _P, _P, _P, _IP, _P],
    "svfsi_lhsa": [_I, _I, _I, _IP, _IP, _IP, _IP],
    "svfsi_mesh_create": [C.POINTER(_P), _P, _I, _I, _I, _IP, _DP],
    "svfsi_mesh_free": [_P],
    "svfsi_mesh_ncolors": [_P],
    "svfsi_mesh_add_face": [_P, _I, _IP, _IP, _IP],
    "svfsi_zero_dev": [_P, _I, _P, _P, _P],
    "svfsi_assemble_dev": [_P, _I, C.POINTER(EqParams), _I, _P, _P, _P, _P, _I, _P, _P, _P],
    "svfsi_assemble_traction_dev": [_P, _I, _I, _DP, _P, _P],
    "svfsi_assemble_backflow_dev": [_P, _I, C.POINTER(EqParams), _I, _P, _P, _P, _P],
    "svfsi_lelas_dkde_dot_dev": [_P, C.POINTER(EqParams), _P, _P, _P, _P],
    "svfsi_fluid_dmu_dot_dev": [_P, C.POINTER(EqParams), _I, _P, _P, _P, _P, _P, _P, _P],
    "svfsi_poly_factors": [_I, _I, _DP, _DP, _IP, _I],
    "svfsi_read_ccne_sizes": [C.c_char_p, C.c_char_p, _IP, _IP, _IP, _IP],
    "svfsi_read_ccne": [C.c_char_p, C.c_char_p, _I, _I, _I, _I, _IP, _DP],
    "svfsi_read_ebc_sizes": [C.c_char_p, _IP, _IP],
    "svfsi_read_ebc": [C.c_char_p, _I, _I, _IP, _IP],
    "svfsi_checkien": [_I, _I, _I, _IP, _DP, _IP],
    "svfsi_read_vtk_sizes": [C.c_char_p, _IP, _IP, _IP, _IP, _IP],
    "svfsi_read_vtk": [C.c_char_p, _I, _I, _I, _DP, _IP, _IP, _IP],
    "svfsi_prof_enable": [_I],
    "svfsi_prof_get": [_I, C.POINTER(C.c_longlong), _DP],
    "svfsi_prof_get_bytes": [_I, C.POINTER(C.c_double)],
    "svfsi_prof_reset": [],
    "svfsi_fsils_fatal": [_I],
    "svfsi_fsils_status": [],
    "svfsi_dev_malloc": [C.POINTER(_P), C.c_size_t],
    "svfsi_dev_free": [_P],
    "svfsi_memcpy_h2d": [_P, _P, C.c_size_t],
    "svfsi_memcpy_d2h": [_P, _P, C.c_size_t],
    "fsils_gpu_solve_cg_noalloc_dev": [_P, C.POINTER(FSILS_lsType), _I, _P, _P, _IP, _P],
}

# The reference's own C API (svFSILS/FSILS_api_c.h:52-64): every argument by
# pointer, no return value; misuse stops the program unless
# svfsi_fsils_fatal(0) was called (then read svfsi_fsils_status()).
VOID_SIGNATURES = {
    "fsils_commu_create_": [C.POINTER(FSILS_commuType), C.POINTER(C.c_ssize_t)],
    "fsils_commu_free_": [C.POINTER(FSILS_commuType)],
    "fsils_lhs_create_c_": [C.POINTER(_P), C.POINTER(FSILS_commuType), _IP, _IP, _IP, _IP, _IP, _IP, _IP],
    "fsils_lhs_free_": [_P],
    "fsils_ls_create_": [C.POINTER(FSILS_lsType), _IP, _DP, _DP, _IP, _IP, _DP, _DP, _IP],
    "fsils_ls_free_": [C.POINTER(FSILS_lsType)],
    "fsils_bc_create_": [_P, _IP, _IP, _IP, _IP, _IP, _DP],
    "fsils_bc_free_": [_P, _IP],
    "fsils_solve_": [_P, C.POINTER(FSILS_lsType), _IP, _DP, _DP, _IP, _IP, _DP],
}


class SvfsiError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load libsvfsi_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SvfsiError(
                f"{LIB_PATH} is missing: build it with `python -m svfsi_amd.build` "
                "(the HIP library is required; there is no CPU fallback)")
        # torch ships its own libamdhip64.so.7.  Loading it first makes the
        # library's NEEDED libamdhip64.so.7 resolve to the same runtime, so
        # torch tensors, streams and this library share one HIP runtime
        # (two runtimes in one process cannot share device pointers).
        import torch  # noqa: F401

        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        for name, args in VOID_SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        L.svfsi_noalloc_scratch_bytes.argtypes = []
        L.svfsi_noalloc_scratch_bytes.restype = C.c_size_t
        L.svfsi_last_error.argtypes = []
        L.svfsi_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != SVFSI_OK:
        msg = lib().svfsi_last_error().decode(errors="replace")
        raise SvfsiError(f"{what} failed (rc={rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
