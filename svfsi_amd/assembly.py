"""Device-resident element assembly over the libsvfsi_amd C ABI.

Replaces the element loop of GLOBALEQASSEM (svFSI/EQASSEM.f:38-87) and the
boundary-face terms (SETBCTRACL, svFSI/SETBC.f:319-416; BASSEMNEUBC + BFLUID,
EQASSEM.f:90-192, FLUID.f:1280-1336).  State and output arrays are torch
CUDA tensors in the reference layout: Ag/Yg/Dg (nNo, tDof), Val (nnz,
dof*dof), R (nNo, dof).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import EqParams, call, lib
from .fsils import Lhs, cur_stream, dev_ptr, _dptr, _i32, _iptr, _f64
from .mesh import TetMesh


def eq_params(d: dict) -> EqParams:
    p = EqParams()
    for k in ("am", "af", "beta", "gam", "dt", "rho", "E", "nu", "mu", "bstab", "C10", "Kpen"):
        setattr(p, k, float(d.get(k, 0.0)))
    fb = d.get("fb", (0.0, 0.0, 0.0))
    for i in range(3):
        p.fb[i] = float(fb[i])
    p.volType = int(d.get("volType", 2))
    # viscosity model (READVISCMODEL): constant unless viscType says otherwise
    p.viscType = int(d.get("viscType", 698))
    for k in ("mu_o", "lam", "visc_a", "visc_n"):
        setattr(p, k, float(d.get(k, 0.0)))
    return p


def _opt(t):
    return None if t is None else dev_ptr(t)


class DeviceMesh:
    """Mesh bound to an lhs pattern: coordinates, colour-sorted connectivity
    and the DOASSEM scatter map, resident on the device."""

    def __init__(self, lhs: Lhs, m: TetMesh):
        self.lhs = lhs
        self.mesh = m
        ien1 = _i32(m.ien) + 1
        x = _f64(m.x)
        h = C.c_void_p()
        call("svfsi_mesh_create", C.byref(h), lhs.handle, m.nNo, m.nEl, m.eNoN, _iptr(ien1), _dptr(x))
        self.handle = h

    @property
    def ncolors(self) -> int:
        return lib().svfsi_mesh_ncolors(self.handle)

    def add_face(self, fien: np.ndarray, fel: np.ndarray) -> int:
        f1 = _i32(fien) + 1
        e1 = _i32(fel) + 1
        fid = C.c_int(-1)
        call("svfsi_mesh_add_face", self.handle, len(f1), _iptr(f1), _iptr(e1), C.byref(fid))
        return fid.value

    def zero(self, dof: int, Val, R, stream=None) -> None:
        call("svfsi_zero_dev", self.handle, dof, dev_ptr(Val), dev_ptr(R), stream or cur_stream())

    def assemble(self, phys: int, eq: dict, Ag, Yg, Dg, Val, R, prop_el=None, quad: int = 0,
                 stream=None) -> None:
        tDof = (Ag if Ag is not None else Yg).shape[1]
        p = eq_params(eq)
        call("svfsi_assemble_dev", self.handle, phys, C.byref(p), tDof, _opt(Ag), _opt(Yg), _opt(Dg),
             _opt(prop_el), quad, dev_ptr(Val), dev_ptr(R), stream or cur_stream())

    def traction(self, face_id: int, dof: int, h, R, stream=None) -> None:
        hh = _f64(h)
        call("svfsi_assemble_traction_dev", self.handle, face_id, dof, _dptr(hh), dev_ptr(R),
             stream or cur_stream())

    def backflow(self, face_id: int, eq: dict, Yg, Val, R, stream=None) -> None:
        p = eq_params(eq)
        call("svfsi_assemble_backflow_dev", self.handle, face_id, C.byref(p), Yg.shape[1], dev_ptr(Yg),
             dev_ptr(Val), dev_ptr(R), stream or cur_stream())

    def lelas_dkde_dot(self, eq: dict, lam, u, g, stream=None) -> None:
        p = eq_params(eq)
        call("svfsi_lelas_dkde_dot_dev", self.handle, C.byref(p), dev_ptr(lam), dev_ptr(u), dev_ptr(g),
             stream or cur_stream())

    def fluid_dmu_dot(self, eq: dict, Ag, Yg, mu_el, lam, du, g, stream=None) -> None:
        """g[e] = lam_e . (dR_e/dmu_e - dK_e/dmu_e du_e) (exact element derivative)."""
        p = eq_params(eq)
        call("svfsi_fluid_dmu_dot_dev", self.handle, C.byref(p), Yg.shape[1], dev_ptr(Ag), dev_ptr(Yg),
             _opt(mu_el), dev_ptr(lam), dev_ptr(du), dev_ptr(g), stream or cur_stream())

    def free(self) -> None:
        if self.handle:
            call("svfsi_mesh_free", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
