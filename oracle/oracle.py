"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY -- the checker for the HIP product path.  The
restated algorithms are in svfsi_oracle.c, each function citing the
reference routine it follows.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "liboracle.so"
SRC = HERE / "svfsi_oracle.c"


def build() -> Path:
    """Compile the oracle with gcc (OpenMP from gcc's own runtime; no other
    libraries)."""
    tmp = SO.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-fopenmp", "-Wall", "-Wextra", "-o", str(tmp), str(SRC),
           "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stderr}")
    tmp.replace(SO)
    return SO


class Eq(C.Structure):
    _fields_ = [
        ("am", C.c_double), ("af", C.c_double), ("beta", C.c_double), ("gam", C.c_double),
        ("dt", C.c_double), ("rho", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("fb", C.c_double * 3), ("mu", C.c_double), ("bstab", C.c_double),
        ("C10", C.c_double), ("Kpen", C.c_double), ("volType", C.c_int), ("pad", C.c_int),
        ("viscType", C.c_int), ("pad2", C.c_int), ("mu_o", C.c_double), ("lam", C.c_double),
        ("visc_a", C.c_double), ("visc_n", C.c_double),
    ]


class SubLs(C.Structure):
    _fields_ = [
        ("suc", C.c_int), ("mItr", C.c_int), ("sD", C.c_int), ("itr", C.c_int),
        ("cM", C.c_int), ("cN", C.c_int), ("cD", C.c_int), ("reserve", C.c_int),
        ("absTol", C.c_double), ("relTol", C.c_double), ("iNorm", C.c_double),
        ("fNorm", C.c_double), ("dB", C.c_double), ("callD", C.c_double),
    ]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        srcs = [SRC, HERE / "svfsi_oracle.h", HERE / "fluid_elem.inc"]
        if not SO.exists() or SO.stat().st_mtime < max(p.stat().st_mtime for p in srcs):
            build()
        _lib = C.CDLL(str(SO))
        d = C.c_double
        _lib.orc_dot.restype = d
        _lib.orc_norm.restype = d
        _lib.orc_iszero.restype = C.c_int
        _lib.orc_lhsa_count.restype = C.c_int
        _lib.orc_fsils_solve.restype = C.c_int
        _lib.orc_diagptr.restype = C.c_int
    return _lib


def eq_struct(d: dict) -> Eq:
    e = Eq()
    for k in ("am", "af", "beta", "gam", "dt", "rho", "E", "nu", "mu", "bstab", "C10", "Kpen"):
        setattr(e, k, float(d.get(k, 0.0)))
    fb = d.get("fb", (0.0, 0.0, 0.0))
    for i in range(3):
        e.fb[i] = float(fb[i])
    e.volType = int(d.get("volType", 2))
    e.viscType = int(d.get("viscType", 698))
    for k in ("mu_o", "lam", "visc_a", "visc_n"):
        setattr(e, k, float(d.get(k, 0.0)))
    return e


def P(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- pattern
def lhsa(nNo: int, ien: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """LHSA restatement: 0-based rowPtr, colPtr."""
    ien = i32(ien)
    nEl, eNoN = ien.shape
    rowPtr = np.zeros(nNo + 1, np.int32)
    nnz = lib().orc_lhsa_count(nNo, nEl, eNoN, P(ien), P(rowPtr))
    if nnz < 0:
        raise ValueError(f"node {-nnz} is isolated")
    colPtr = np.zeros(nnz, np.int32)
    lib().orc_lhsa_fill(nNo, nEl, eNoN, P(ien), P(rowPtr), P(colPtr))
    return rowPtr, colPtr


# ---------------------------------------------------------------- assembly
def assemble(prob, rowPtr: np.ndarray, colPtr: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Element loop + DOASSEM + face terms of a Problem (0-based pattern)."""
    m = prob.mesh
    dof = prob.dof
    Val = np.zeros((len(colPtr), dof * dof))
    R = np.zeros((m.nNo, dof))
    eq = eq_struct(prob.eq)
    x, ien = f64(m.x), i32(m.ien)
    Ag, Yg, Dg = f64(prob.Ag), f64(prob.Yg), f64(prob.Dg)
    prop = None if prob.prop_el is None else P(f64(prob.prop_el))
    L = lib()
    if prob.phys == 2:
        L.orc_construct_lelas(m.nNo, m.nEl, P(ien), P(x), P(Ag), P(Dg), Ag.shape[1], C.byref(eq), prop,
                              P(rowPtr), P(colPtr), P(Val), P(R))
    elif prob.phys == 1 and m.eNoN == 10:
        # Taylor-Hood (nFs = 2): element loop, then THOOD_ValRC after the
        # face terms, where svFSI's MAIN calls it (MAIN.f:167-173)
        if L.orc_construct_fluid_th(m.nNo, m.nEl, P(ien), P(x), P(Ag), P(Yg), Ag.shape[1], C.byref(eq), prop,
                                    P(rowPtr), P(colPtr), P(Val), P(R)):
            raise RuntimeError("GNNxx: singular element")
    elif prob.phys == 1:
        L.orc_construct_fluid(m.nNo, m.nEl, P(ien), P(x), P(Ag), P(Yg), Ag.shape[1], C.byref(eq), prop,
                              P(rowPtr), P(colPtr), P(Val), P(R))
    elif prob.phys == 3:
        L.orc_construct_struct(m.nNo, m.nEl, m.eNoN, P(ien), P(x), P(Ag), P(Yg), P(Dg), Ag.shape[1],
                               C.byref(eq), prob.quad & 1, P(rowPtr), P(colPtr), P(Val), P(R))
    for fien, fel, h in prob.trac_faces:
        fi, fe, hh = i32(fien), i32(fel), f64(h)
        L.orc_traction_tri3(len(fi), P(fi), P(fe), P(ien), m.eNoN, P(x), P(hh), dof, P(rowPtr), P(colPtr),
                            P(Val), P(R))
    for fien, fel in prob.bflow_faces:
        fi, fe = i32(fien), i32(fel)
        L.orc_backflow_tri3(len(fi), P(fi), P(fe), P(ien), m.eNoN, P(x), P(Yg), Yg.shape[1], C.byref(eq),
                            P(rowPtr), P(colPtr), P(Val), P(R))
    if prob.phys == 1 and m.eNoN == 10:
        L.orc_thood_valrc(m.nNo, m.nEl, P(ien), P(rowPtr), P(colPtr), P(Val), P(R))
    return Val, R


def gnnxx(Nxi: np.ndarray, Nxi2: np.ndarray, xl: np.ndarray, Nx: np.ndarray) -> np.ndarray:
    """GNNxx (svFSI/NN.f:1568-1678): physical second derivatives (eNoN, 6)
    in the order xx, yy, zz, xy, yz, xz."""
    Nxx = np.zeros((len(Nxi), 6))
    if lib().orc_gnnxx(len(Nxi), P(f64(Nxi)), P(f64(Nxi2)), P(f64(xl)), P(f64(Nx)), P(Nxx)):
        raise RuntimeError("GNNxx: singular system")
    return Nxx


def shape_tet10(xi) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(N (10), Nxi (10, 3), Nxi2 (10, 6)) of the tet10 basis at xi."""
    N = np.zeros(10)
    Nxi = np.zeros((10, 3))
    Nxi2 = np.zeros((10, 6))
    lib().orc_shape_tet10(P(f64(xi)), P(N), P(Nxi))
    lib().orc_shape_tet10_xx(P(Nxi2))
    return N, Nxi, Nxi2


def getviscosity(eq: dict, gamma: float, mu_i: float | None = None):
    """GETVISCOSITY at shear rate gamma: (mu, mu_s, d mu / d gamma, gamma
    after the call: Casson raises it to its threshold)."""
    e = eq_struct(eq)
    g, mu, mus, mux = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    fn = lib().orc_getviscosity
    fn.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    fn(C.byref(e), float(eq.get("mu", 0.0) if mu_i is None else mu_i), float(gamma), C.byref(g), C.byref(mu),
       C.byref(mus), C.byref(mux))
    return mu.value, mus.value, mux.value, g.value


def set_threads(n: int) -> None:
    """Threads of the oracle's solver kernels (CPU baseline timing only)."""
    lib().orc_set_threads(int(n))


def assemble_ranks_seconds(prob, rowPtr: np.ndarray, colPtr: np.ndarray, nranks: int) -> float:
    """Wall time of the element loop split over nranks host threads, each
    assembling a contiguous range of elements into its own zeroed Val and R
    (the per-rank local systems of the reference's MPI partition; the
    interface sum that follows there is a halo exchange and is not
    emulated; the outputs are zeroed beforehand, as each rank's share of
    LSALLOC's zeroing is 1/nranks of the matrix).  The ctypes calls release the GIL, so the ranks run
    concurrently.  Face terms are left out (a few hundred elements)."""
    import concurrent.futures as cf
    import time

    m = prob.mesh
    dof = prob.dof
    eq = eq_struct(prob.eq)
    x, ien = f64(m.x), i32(m.ien)
    Ag, Yg, Dg = f64(prob.Ag), f64(prob.Yg), f64(prob.Dg)
    prop = None if prob.prop_el is None else f64(prob.prop_el)
    rp, cp = i32(rowPtr), i32(colPtr)
    bounds = [int(b) for b in np.linspace(0, m.nEl, nranks + 1)]
    # per-rank inputs and zeroed outputs, prepared outside the timed region
    bufs = [(np.zeros((len(colPtr), dof * dof)), np.zeros((m.nNo, dof))) for _ in range(nranks)]
    subs = [np.ascontiguousarray(ien[bounds[k]:bounds[k + 1]]) for k in range(nranks)]
    props = [None if prop is None else np.ascontiguousarray(prop[bounds[k]:bounds[k + 1]]) for k in range(nranks)]
    L = lib()

    def rank(k):
        e0, e1 = bounds[k], bounds[k + 1]
        Val, R = bufs[k]
        sub = subs[k]
        pr = None if props[k] is None else P(props[k])
        if prob.phys == 2:
            L.orc_construct_lelas(m.nNo, e1 - e0, P(sub), P(x), P(Ag), P(Dg), Ag.shape[1], C.byref(eq), pr,
                                  P(rp), P(cp), P(Val), P(R))
        elif prob.phys == 1:
            L.orc_construct_fluid(m.nNo, e1 - e0, P(sub), P(x), P(Ag), P(Yg), Ag.shape[1], C.byref(eq), pr,
                                  P(rp), P(cp), P(Val), P(R))
        else:
            L.orc_construct_struct(m.nNo, e1 - e0, m.eNoN, P(sub), P(x), P(Ag), P(Yg), P(Dg), Ag.shape[1],
                                   C.byref(eq), prob.quad & 1, P(rp), P(cp), P(Val), P(R))

    with cf.ThreadPoolExecutor(max_workers=nranks) as ex:
        list(ex.map(lambda k: None, range(nranks)))  # start the workers
        t0 = time.perf_counter()
        list(ex.map(rank, range(nranks)))
        t1 = time.perf_counter()
    return t1 - t0


# ---------------------------------------------------------------- solver
def spmv(rowPtr, colPtr, dof, K, U) -> np.ndarray:
    nNo = len(rowPtr) - 1
    out = np.zeros((nNo, dof))
    lib().orc_spmv(nNo, dof, P(i32(rowPtr)), P(i32(colPtr)), P(f64(K)), P(f64(U)), P(out))
    return out


def spmv_t(rowPtr, colPtr, dof, K, U) -> np.ndarray:
    nNo = len(rowPtr) - 1
    out = np.zeros((nNo, dof))
    lib().orc_spmv_t(nNo, dof, P(i32(rowPtr)), P(i32(colPtr)), P(f64(K)), P(f64(U)), P(out))
    return out


def precond_diag(rowPtr, colPtr, dof, Val, R, mask):
    """Returns (scaled Val, scaled R, W)."""
    nNo = len(rowPtr) - 1
    rp, cp = i32(rowPtr), i32(colPtr)
    dg = np.zeros(nNo, np.int32)
    assert lib().orc_diagptr(nNo, P(rp), P(cp), P(dg)) == 0
    V, Rr = f64(Val).copy(), f64(R).copy()
    W = np.zeros((nNo, dof))
    lib().orc_precond_diag(nNo, dof, P(rp), P(cp), P(dg), None if mask is None else P(f64(mask)), P(V), P(Rr), P(W))
    return V, Rr, W


def _ls(settings: dict) -> SubLs:
    s = SubLs()
    s.relTol = settings["relTol"]
    s.absTol = settings["absTol"]
    s.mItr = settings["mItr"]
    s.sD = settings.get("sD", 0) or 0
    return s


def cgradv(rowPtr, colPtr, dof, K, R, settings: dict):
    x = f64(R).copy()
    s = _ls(settings)
    lib().orc_cgradv(len(rowPtr) - 1, dof, P(i32(rowPtr)), P(i32(colPtr)), P(f64(K)), P(x), C.byref(s))
    return x, s


def gmresv(rowPtr, colPtr, dof, K, R, settings: dict):
    x = f64(R).copy()
    s = _ls(settings)
    lib().orc_gmresv(len(rowPtr) - 1, dof, P(i32(rowPtr)), P(i32(colPtr)), P(f64(K)), P(x), C.byref(s))
    return x, s


def precond_rcs(rowPtr, colPtr, dof, Val, R, mask):
    """PRECONDRCS.  Returns (scaled Val, scaled R, W1, W2)."""
    nNo = len(rowPtr) - 1
    rp, cp = i32(rowPtr), i32(colPtr)
    dg = np.zeros(nNo, np.int32)
    assert lib().orc_diagptr(nNo, P(rp), P(cp), P(dg)) == 0
    V, Rr = f64(Val).copy(), f64(R).copy()
    W1 = np.zeros((nNo, dof))
    W2 = np.zeros((nNo, dof))
    lib().orc_precond_rcs(nNo, dof, P(rp), P(cp), P(dg), None if mask is None else P(f64(mask)), P(V), P(Rr),
                          P(W1), P(W2))
    return V, Rr, W1, W2


def fsils_solve(rowPtr, colPtr, dof, mask, settings: dict, R, Val, prec: int = 701):
    """FSILS_SOLVE restatement (prec 701 PRECOND_FSILS, 709 PRECOND_RCS).
    Returns (x, scaled Val, SubLs)."""
    x = f64(R).copy()
    V = f64(Val).copy()
    s = _ls(settings)
    rc = lib().orc_fsils_solve_prec(len(rowPtr) - 1, dof, P(i32(rowPtr)), P(i32(colPtr)),
                                    None if mask is None else P(f64(mask)), settings["LS_type"], int(prec),
                                    C.byref(s), P(x), P(V))
    if rc != 0:
        raise RuntimeError(f"orc_fsils_solve rc={rc}")
    return x, V, s


class Ls(C.Structure):
    """FSILS_lsType: the outer record RI and the NS solver's inner GM, CG."""
    _fields_ = [("foC", C.c_int), ("LS_type", C.c_int), ("Resm", C.c_int), ("Resc", C.c_int),
                ("GM", SubLs), ("CG", SubLs), ("RI", SubLs)]


def ns_ls(settings: dict) -> Ls:
    """NS solver record: RI from settings (relTol, absTol, mItr), GM and CG
    from settings["GM"] / settings["CG"] (relTol, absTol, mItr, sD), with the
    reference's NS defaults (LS.f:69-95) for what is missing."""
    ls = Ls()
    ls.LS_type = 796
    ls.RI.relTol = settings.get("relTol", 0.4)
    ls.RI.absTol = settings.get("absTol", 1e-10)
    ls.RI.mItr = settings.get("mItr", 10)
    ls.RI.sD = settings.get("sD") or 100
    gm = {**dict(relTol=1e-2, absTol=1e-10, mItr=2, sD=ls.RI.sD), **settings.get("GM", {})}
    cg = {**dict(relTol=0.2, absTol=1e-10, mItr=500), **settings.get("CG", {})}
    for rec, d in ((ls.GM, gm), (ls.CG, cg)):
        rec.relTol, rec.absTol, rec.mItr = d["relTol"], d["absTol"], d["mItr"]
        rec.sD = d.get("sD", 0) or 0
    return ls


def fsils_solve_ns(rowPtr, colPtr, dof, mask, settings: dict, R, Val, prec: int = 701):
    """FSILS_SOLVE with the NS solver (NSSOLVER).  Returns (x, scaled Val, Ls)."""
    x = f64(R).copy()
    V = f64(Val).copy()
    ls = ns_ls(settings)
    rc = lib().orc_fsils_solve_ls(len(rowPtr) - 1, dof, P(i32(rowPtr)), P(i32(colPtr)),
                                  None if mask is None else P(f64(mask)), 796, int(prec), C.byref(ls), P(x), P(V))
    if rc != 0:
        raise RuntimeError(f"orc_fsils_solve_ls rc={rc}")
    return x, V, ls


def coupled_faces(faces, dof: int):
    """Concatenated coupled Neumann faces for orc_fsils_solve_cpl: faces is a
    list of (nodes 0-based, face dof, val (nNo_f, face dof), res)."""
    off = [0]
    for nodes, _, _, _ in faces:
        off.append(off[-1] + len(nodes))
    glob = np.concatenate([np.asarray(nd, dtype=np.int32) for nd, _, _, _ in faces])
    fdof = np.array([fd for _, fd, _, _ in faces], dtype=np.int32)
    val = np.zeros((off[-1], dof))
    for f, (nodes, fd, v, _) in enumerate(faces):
        k = min(fd, dof)
        val[off[f]:off[f + 1], :k] = np.asarray(v, dtype=float).reshape(len(nodes), fd)[:, :k]
    res = np.array([r for _, _, _, r in faces], dtype=float)
    return i32(off), glob, fdof, val, res


def fsils_solve_cpl(rowPtr, colPtr, dof, mask, settings: dict, R, Val, faces, prec: int = 701):
    """FSILS_SOLVE with coupled Neumann faces (ADDBCMUL in GMRESV / NSSOLVER;
    CGRADV ignores them).  settings["LS_type"] 796, 797 or 798.  Returns
    (x, scaled Val, Ls): Ls.RI is the outer record for every solver."""
    x = f64(R).copy()
    V = f64(Val).copy()
    lt = settings["LS_type"]
    if lt == 796:
        ls = ns_ls(settings)
    else:
        ls = Ls()
        ls.LS_type = lt
        ls.RI = _ls(settings)
    off, glob, fdof, val, res = coupled_faces(faces, dof)
    rc = lib().orc_fsils_solve_cpl(len(rowPtr) - 1, dof, P(i32(rowPtr)), P(i32(colPtr)),
                                   None if mask is None else P(f64(mask)), lt, int(prec), C.byref(ls),
                                   len(faces), P(off), P(glob), P(fdof), P(f64(val)), P(f64(res)), P(x), P(V))
    if rc != 0:
        raise RuntimeError(f"orc_fsils_solve_cpl rc={rc}")
    return x, V, ls


def fluid_dmu_dot(prob, lam, du) -> np.ndarray:
    """g[e] = lam_e . (dR_e/dmu_e - dK_e/dmu_e du_e) (complex step)."""
    m = prob.mesh
    g = np.zeros(m.nEl)
    eq = eq_struct(prob.eq)
    mu = None if prob.prop_el is None else P(f64(prob.prop_el))
    Ag, Yg = f64(prob.Ag), f64(prob.Yg)
    lib().orc_fluid_dmu_dot(m.nEl, P(i32(m.ien)), P(f64(m.x)), P(Ag), P(Yg), Ag.shape[1], C.byref(eq), mu,
                            P(f64(lam)), P(f64(du)), P(g))
    return g


def lelas_dkde_dot(prob, lam, u) -> np.ndarray:
    m = prob.mesh
    g = np.zeros(m.nEl)
    eq = eq_struct(prob.eq)
    lib().orc_lelas_dkde_dot(m.nEl, P(i32(m.ien)), P(f64(m.x)), C.byref(eq), None, P(f64(lam)), P(f64(u)), P(g))
    return g
