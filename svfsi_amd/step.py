"""One Newton step of the linear-system path on the GPU.

Mirrors the sequence of one svFSI Newton iteration between PICI and PICC
(svFSI/MAIN.f:116-209): LSALLOC zeroes R and Val (svFSI/LS.f:338-366),
GLOBALEQASSEM assembles the element terms (EQASSEM.f:38-87), SETBCNEU adds
the face terms (SETBC.f:231-249), LSSOLVE calls FSILS_SOLVE (LS.f:368-410).
Everything after set-up stays resident in HBM.
"""
from __future__ import annotations

import numpy as np

from . import fsils
from ._lib import BC_TYPE_Dir, BC_TYPE_Neu, SVFSI_FLUID_P2P1, FSILS_lsType
from .assembly import DeviceMesh
from .problems import Problem


class NewtonStep:
    def __init__(self, prob: Problem):
        import torch

        self.prob = prob
        m = prob.mesh
        self.rowPtr, self.colPtr = fsils.lhsa(m.nNo, m.ien)
        self.nnz = int(self.colPtr.shape[0])
        nDir = len(prob.dir_faces)
        self.lhs = fsils.fsils_lhs_create(m.nNo, self.nnz, self.rowPtr, self.colPtr,
                                          nFaces=nDir + len(prob.res_faces))
        for i, (nodes, fdof) in enumerate(prob.dir_faces):
            fsils.fsils_bc_create(self.lhs, i + 1, len(nodes), fdof, BC_TYPE_Dir, np.asarray(nodes) + 1)
        # resistance faces (svFSI/BAFINI.f:512-533): Neumann faces whose face
        # vector couples through res in FSILS_SOLVE
        for i, (nodes, fdof, val, _) in enumerate(prob.res_faces):
            fsils.fsils_bc_create(self.lhs, nDir + i + 1, len(nodes), fdof, BC_TYPE_Neu, np.asarray(nodes) + 1,
                                  np.asarray(val, dtype=np.float64).reshape(len(nodes), fdof))
        self.res = None
        if prob.res_faces:
            self.res = np.array([0.0] * nDir + [float(r) for _, _, _, r in prob.res_faces])
        self.dmesh = DeviceMesh(self.lhs, m)
        self.trac_ids = [(self.dmesh.add_face(f, e), h) for f, e, h in prob.trac_faces]
        self.bflow_ids = [self.dmesh.add_face(f, e) for f, e in prob.bflow_faces]
        dev = torch.device("cuda")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
        self.Ag, self.Yg, self.Dg = t(prob.Ag), t(prob.Yg), t(prob.Dg)
        self.prop = None if prob.prop_el is None else t(prob.prop_el)
        dof = prob.dof
        self.Val = torch.zeros((self.nnz, dof * dof), dtype=torch.float64, device=dev)
        self.R = torch.zeros((m.nNo, dof), dtype=torch.float64, device=dev)

    @property
    def ncolors(self) -> int:
        return self.dmesh.ncolors

    def assemble(self) -> None:
        p = self.prob
        self.dmesh.zero(p.dof, self.Val, self.R)
        quad = SVFSI_FLUID_P2P1 if p.nFs == 2 else p.quad
        self.dmesh.assemble(p.phys, p.eq, self.Ag, self.Yg, self.Dg, self.Val, self.R,
                            prop_el=self.prop, quad=quad)
        for fid, h in self.trac_ids:
            self.dmesh.traction(fid, p.dof, h, self.R)
        for fid in self.bflow_ids:
            self.dmesh.backflow(fid, p.eq, self.Yg, self.Val, self.R)

    def make_ls(self) -> FSILS_lsType:
        c = self.prob.ls
        inner = {}
        if c["LS_type"] == fsils.LS_TYPE_NS:
            # the NS solver's momentum GMRES and Schur-complement CG records
            # (LS.f:69-95 defaults where the settings leave them out)
            gm = {**dict(relTol=1e-2, absTol=1e-10, mItr=2), **c.get("GM", {})}
            cg = {**dict(relTol=0.2, absTol=1e-10, mItr=500), **c.get("CG", {})}
            inner = dict(relTolIn=[gm["relTol"], cg["relTol"]], absTolIn=[gm["absTol"], cg["absTol"]],
                         maxItrIn=[gm["mItr"], cg["mItr"]])
        return fsils.fsils_ls_create(c["LS_type"], relTol=c["relTol"], absTol=c["absTol"],
                                     maxItr=c["mItr"], dimKry=c["sD"] if c.get("sD") else None, **inner)

    def solve(self, mixed: bool = False) -> FSILS_lsType:
        ls = self.make_ls()
        prec = self.prob.ls.get("prec", fsils.PRECOND_FSILS)
        fsils.fsils_solve(self.lhs, ls, self.prob.dof, self.R, self.Val, prec=prec, res=self.res, mixed=mixed)
        return ls

    def step(self, mixed: bool = False) -> FSILS_lsType:
        self.assemble()
        return self.solve(mixed=mixed)
