#!/usr/bin/env python3
"""Benchmark: assembly + GMRES time per Newton step to relative residual
1e-8, FP64 vs mixed precision, on one MI355X.

Workload (default): incompressible Navier-Stokes VMS (FLUID3D_M/C), tet4,
straight cylinder r = 1, L = 10, ~200k tets from a seeded Delaunay
tetrahedralisation of jittered lattice points; rho = 1.06, mu = 0.04,
parabolic inlet peak 30, zero-traction outlet with backflow stabilisation,
no-slip wall, generalized-alpha rho_inf = 0.2, dt = 1e-3, random initial
velocity U(-1, 1); one Newton step = zero + element assembly + face terms +
FSILS_SOLVE with GMRES(50) and diagonal scaling to 1e-8 in the reference's
scaled norm.  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  `value` is the mixed-precision path's time per
Newton step (ms, lower is better); the FP64 path (GMRESV), the same nested
solver with its inner preconditioner in FP64, the SpMV roofline and the
CPU baseline (the FP64 C oracle on this host's cores, one thread per core
standing in for the reference's MPI ranks; one core beside it) are reported
beside it.  `--ls ns` measures the same Newton step with the NS solver
(FP64, the reference's NSSOLVER defaults) as a separate line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"mfma64": 78.6, "mfma32": 157.3}  # dense FP64 / FP32 matrix peaks


def spmv_bytes(nnzb: int, nNo: int, dof: int, s: int) -> int:
    """Algorithmic bytes of one block SpMV: values, column indices, row
    pointers, one read of U and one write of KU."""
    return nnzb * dof * dof * s + 4 * nnzb + 4 * (nNo + 1) + 2 * nNo * dof * s


def log(*a) -> None:
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print("#", *a, file=sys.stderr, flush=True)


def measured_traffic(nnzb: int, kernel: str):
    """HBM bytes per launch of `kernel` from the PMC summary in profiles/
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy kernel of known
    size), when it was measured on a matrix of the same size; else None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel", "").startswith(kernel.split("*")[0]) and r.get("nnzb") == nnzb:
            return r.get("traffic_bytes")
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["fluid", "fluid-large", "struct", "lelas", "fluid-th"], default="fluid",
                    help="fluid: VMS cylinder ~200k tet4 (the metric's configuration); fluid-large: the same physics "
                         "on ~3.2M tet4 with GMRES(100); struct: neo-Hookean tet10 cube 40^3 x 6, 4-point rule; "
                         "lelas: linear elasticity cube 20^3 x 6 tet4 with CG; fluid-th: Taylor-Hood (P2P1) "
                         "channel on the tet10 cube 20^3 x 6 with GMRES(50)")
    ap.add_argument("--n", type=int, default=0, help="cube cells per side for struct/lelas (default 40/20)")
    ap.add_argument("--h", type=float, default=0.11, help="cylinder lattice spacing (0.11 -> ~200k tets)")
    ap.add_argument("--sD", type=int, default=0, help="GMRES restart length (default: 50 fluid, 100 struct)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds of all-core CPU solve the baseline may take in full (beyond it, a bounded "
                         "sample is scaled)")
    ap.add_argument("--mitr", type=int, default=0,
                    help="cap the solver's restart count (profiling and sweeps; 0 keeps the workload's setting)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (calibrates PMC byte counters)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="skip the FP64 leg (for counter-collection runs, which serialise every dispatch)")
    ap.add_argument("--ls", choices=["gmres", "ns"], default="gmres",
                    help="ns: the fluid workloads with the NS solver (NSSOLVER, reference defaults) in place of "
                         "GMRES; a separate measurement line (FP64 only, as the reference's NS solver)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU partitioning is a design note, not an implemented path", file=sys.stderr)
        return 2

    import torch

    from svfsi_amd import problems
    from svfsi_amd._lib import SVFSI_STRUCT_MFMA32, SVFSI_STRUCT_MFMA64, SVFSI_STRUCT_VALU, lib
    from svfsi_amd.step import NewtonStep

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1

    if args.copy_probe:
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "reps": 5}))
        return 0

    t0 = time.time()
    if not args.sD:
        args.sD = 100 if args.workload in ("struct", "fluid-large") else 50
    if args.workload in ("fluid", "fluid-large"):
        h = 0.042 if args.workload == "fluid-large" else args.h
        prob = problems.fluid_cylinder(h=h, sD=args.sD)
        if args.workload == "fluid-large":
            prob.ls = dict(prob.ls, mItr=2000)
        wl = "FLUID3D VMS tet4 cylinder r=1 L=10 (h=%g), one Newton step, GMRES(%d) to 1e-8" % (h, args.sD)
    elif args.workload == "fluid-th":
        n = args.n or 20
        prob = problems.fluid_th_cube(n, sD=args.sD)
        wl = f"FLUID Taylor-Hood (P2P1) channel, tet10 cube {n}^3 x 6, one Newton step, GMRES({args.sD}) to 1e-8"
    elif args.workload == "struct":
        n = args.n or 40
        prob = problems.struct_cube(n, quad=1, sD=args.sD)
        wl = f"STRUCT3D neo-Hookean tet10 cube {n}^3 x 6, 4-point rule, one Newton step, GMRES({args.sD}) to 1e-8"
    else:
        n = args.n or 20
        prob = problems.lelas_cube(n)
        wl = f"LELAS3D tet4 cube {n}^3 x 6, one assembly + CG to 1e-8"
    if args.mitr:
        prob.ls = dict(prob.ls, mItr=args.mitr)
    if args.ls == "ns":
        if prob.dof != 4:
            print("bench.py: --ls ns needs a fluid workload (dof 4)", file=sys.stderr)
            return 2
        return bench_ns(args, prob, t0)
    st = NewtonStep(prob)
    t_setup = time.time() - t0
    m = prob.mesh
    nnzb = st.nnz
    dof = prob.dof
    print(f"# setup {t_setup:.1f}s nNo={m.nNo} nEl={m.nEl} nnzb={nnzb} colors={st.ncolors}", file=sys.stderr,
          flush=True)

    def run(mixed: bool):
        ls = st.step(mixed=mixed)
        return ls

    # warmup both paths
    for w in range(args.warmup):
        if not args.mixed_only:
            ls = run(False)
            log(f"warmup {w} fp64: itr {ls.RI.itr} suc {ls.RI.suc}")
        ls = run(True)
        log(f"warmup {w} mixed: outer {ls.RI.itr} spmv {ls.RI.cM} suc {ls.RI.suc}")
    torch.cuda.synchronize()

    def timed(mixed: bool, k: int, profile: bool):
        """Wall time per step over k steps, bracketed by device
        synchronisation.  With profile=True the SpMV launches carry
        dispatch-attached events (kernel durations); such runs are kept out
        of the timed figures."""
        import ctypes as C

        L = lib()
        L.svfsi_prof_reset()
        L.svfsi_prof_enable(1 if profile else 0)
        torch.cuda.synchronize()
        a = time.perf_counter()
        lss = [run(mixed) for _ in range(k)]
        torch.cuda.synchronize()
        b = time.perf_counter()
        L.svfsi_prof_enable(0)
        cnt = [C.c_longlong(), C.c_longlong()]
        tot = [C.c_double(), C.c_double()]
        byt = [C.c_double(), C.c_double()]
        for kk in (0, 1):
            L.svfsi_prof_get(kk, C.byref(cnt[kk]), C.byref(tot[kk]))
            L.svfsi_prof_get_bytes(kk, C.byref(byt[kk]))
        return (b - a) * 1e3 / k, lss, [(cnt[q].value, tot[q].value, byt[q].value) for q in (0, 1)]

    # phase split (outside the timed loops): assembly alone
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(max(1, args.steps)):
        st.assemble()
    torch.cuda.synchronize()
    asm_ms = (time.perf_counter() - a) * 1e3 / max(1, args.steps)

    log(f"assembly {asm_ms:.3f} ms")
    # STRUCT: each element kernel on its own (vector units, FP64 MFMA, FP32
    # MFMA), with the contraction kernels' event timing and MFMA rate
    asm_kern = {}
    struct_kernels = (("valu", SVFSI_STRUCT_VALU), ("mfma64", SVFSI_STRUCT_MFMA64),
                      ("mfma32", SVFSI_STRUCT_MFMA32))
    base_quad = prob.quad
    if args.workload == "struct":
        import ctypes as C

        L = lib()
        for name, kern in struct_kernels:
            prob.quad = (base_quad & 1) | kern
            st.assemble()
            torch.cuda.synchronize()
            a = time.perf_counter()
            for _ in range(max(1, args.steps)):
                st.assemble()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - a) * 1e3 / max(1, args.steps)
            L.svfsi_prof_reset()
            L.svfsi_prof_enable(1)
            st.assemble()
            torch.cuda.synchronize()
            L.svfsi_prof_enable(0)
            cnt, tms, fl = C.c_longlong(), C.c_double(), C.c_double()
            L.svfsi_prof_get(2, C.byref(cnt), C.byref(tms))
            L.svfsi_prof_get_bytes(2, C.byref(fl))
            pc, pms = C.c_longlong(), C.c_double()
            L.svfsi_prof_get(3, C.byref(pc), C.byref(pms))
            asm_kern[name] = {"ms": round(ms, 3), "element_kernels_ms": round(tms.value, 3),
                              "element_launches": cnt.value, "points_kernel_ms": round(pms.value, 3),
                              "contraction_gflop": round(fl.value / 1e9, 2)}
            if name != "valu" and tms.value > 0:
                tf = fl.value / (tms.value * 1e-3) / 1e12
                asm_kern[name]["mfma_tflops"] = round(tf, 2)
                asm_kern[name]["mfma_frac"] = round(tf / MFMA_PEAK_TFLOPS[name], 4)
            log(f"assembly[{name}] {ms:.3f} ms (element kernels {tms.value:.3f} ms, points {pms.value:.3f} ms)")
        prob.quad = base_quad
    if args.mixed_only:
        fp64_ms, ls64, prof64 = None, None, [(0, 0.0, 0.0), (0, 0.0, 0.0)]
    else:
        fp64_ms, ls64, _ = timed(False, args.steps, False)
        log(f"fp64 {fp64_ms:.3f} ms/step")
        _, _, prof64 = timed(False, 1, True)
        log("fp64 profiled run done")
    mix_ms, lsmx, _ = timed(True, args.steps, False)
    log(f"mixed {mix_ms:.3f} ms/step")
    # dof-4 systems whose FP32 matrix exceeds half the Infinity Cache get the
    # multigrid inner preconditioner by default (FP32 only); the polynomial
    # one is then timed beside it, so that the FP64-inner leg below still has
    # its FP32 counterpart
    amg_env = os.environ.get("SVFSI_INNER_AMG", "")
    amg_default = amg_env == "1" or (amg_env != "0" and dof == 4 and nnzb * 64 > (128 << 20))
    poly_ms, lsp = None, None
    if amg_default and not args.mixed_only:
        os.environ["SVFSI_INNER_AMG"] = "0"
        try:
            run(True)
            poly_ms, lsp, _ = timed(True, args.steps, False)
        finally:
            os.environ["SVFSI_INNER_AMG"] = amg_env
            if not amg_env:
                os.environ.pop("SVFSI_INNER_AMG")
        log(f"mixed, polynomial inner {poly_ms:.3f} ms/step (outer {lsp[-1].RI.itr}, SpMVs {lsp[-1].RI.cM})")
    # the same nested solver with its inner preconditioner in FP64: what the
    # FP32 inner storage gains over the algorithm alone
    nest_ms, lsn = None, None
    if not args.mixed_only:
        os.environ["SVFSI_INNER_FP64"] = "1"
        try:
            run(True)
            nest_ms, lsn, _ = timed(True, args.steps, False)
        finally:
            os.environ.pop("SVFSI_INNER_FP64", None)
        log(f"nested fp64 {nest_ms:.3f} ms/step (outer {lsn[-1].RI.itr}, SpMVs {lsn[-1].RI.cM})")
    _, lsprof, profmx = timed(True, 1, True)
    log("mixed profiled run done")

    # SpMV roofline: FP32 SpMV of the mixed path (the inner solver's kernel).
    # Launches queued after an inner cycle has converged exit at once; the
    # achieved rate charges their time to the launches that did the work
    # (inner SpMVs = cM minus the outer FP64 products).
    # The library accumulates each profiled launch's algorithmic bytes (for
    # the fused inner step: SpMV + the basis vectors its multi-dot reads).
    n32, t32, by32 = profmx[1]
    n64m = profmx[0][0]
    n32_full = max(1, min(n32, lsprof[-1].RI.cM - n64m)) if n32 else 0
    if amg_default:
        n32_full = n32  # the V-cycle's launches all do their work (no early exit)
    n64, t64, by64 = prof64[0]
    b32 = by32 / n32 if n32 else spmv_bytes(nnzb, m.nNo, dof, 4)
    b64 = spmv_bytes(nnzb, m.nNo, dof, 8)
    gbs32 = b32 * n32_full / (t32 * 1e-3) / 1e9 if n32 else None
    gbs64 = b64 / (t64 / n64 * 1e-3) / 1e9 if n64 else None

    # accuracy and CPU baseline (outside the timed regions)
    cpu = None
    acc = {}
    if not args.no_cpu and not args.mixed_only:
        from oracle import oracle as orc

        st.assemble()
        torch.cuda.synchronize()
        V0 = st.Val.cpu().numpy().copy()
        R0 = st.R.cpu().numpy().copy()
        mask = prob.dir_mask()
        # CPU leg, bounded sample: the full FP64 oracle assembly, its diagonal
        # scaling, and `cycles` restart cycles of GMRESV; the per-iteration
        # time is scaled to the FP64 iteration count of the same step.
        # host cores: one oracle "rank" (thread) per core of this process's CPU
        # share, as the reference runs one MPI rank per core
        cores = len(os.sched_getaffinity(0))
        cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
        log("cpu leg: oracle assembly (1 core, also the accuracy reference)")
        c0 = time.perf_counter()
        Vo, Ro = orc.assemble(prob, st.rowPtr - 1, st.colPtr - 1)
        c1 = time.perf_counter()
        asm1 = c1 - c0
        asm_par = orc.assemble_ranks_seconds(prob, st.rowPtr - 1, st.colPtr - 1, cores)
        log(f"cpu leg: assembly {c1 - c0:.1f} s on 1 core, {asm_par:.2f} s on {cores}; solver sample")
        orc.set_threads(cores)
        c1 = time.perf_counter()
        Vs, Rs, W = orc.precond_diag(st.rowPtr - 1, st.colPtr - 1, dof, Vo, Ro, mask)
        c2 = time.perf_counter()
        orc.set_threads(1)
        smp = dict(prob.ls)
        if args.workload == "lelas":
            smp["mItr"] = 100
            solve = lambda: orc.cgradv(st.rowPtr - 1, st.colPtr - 1, dof, Vs, Rs, smp)
            what = "CGRADV iterations"
        else:
            cycles = 1 if args.workload == "struct" else 2
            smp["mItr"] = cycles
            solve = lambda: orc.gmresv(st.rowPtr - 1, st.colPtr - 1, dof, Vs, Rs, smp)
            what = f"GMRESV iterations ({cycles} restart cycles)"
        # one core: a bounded sample, scaled to the FP64 iteration count
        orc.set_threads(1)
        a = time.perf_counter()
        _, so1 = solve()
        per_it1 = (time.perf_counter() - a) / max(so1.itr, 1)
        # all cores: the whole FP64 solve, PRECONDDIAG included, not scaled
        # (the bounded sample with `cores` threads only when the full solve
        # would exceed the CPU budget of this run)
        itr64 = ls64[-1].RI.itr
        orc.set_threads(cores)
        a = time.perf_counter()
        _, so = solve()
        per_itc = (time.perf_counter() - a) / max(so.itr, 1)
        full = per_itc * itr64 < args.cpu_budget
        if full:
            full_ls = dict(prob.ls)
            V1, R1 = Vo.copy(), Ro.copy()
            a = time.perf_counter()
            _, _, sfull = orc.fsils_solve(st.rowPtr - 1, st.colPtr - 1, dof, mask, full_ls, R1, V1)
            solve_full = time.perf_counter() - a
            log(f"cpu leg: full FP64 solve on {cores} cores: {solve_full:.1f} s, {sfull.itr} iterations")
        orc.set_threads(1)
        c3 = time.perf_counter()
        if full:
            cpu_ms = (asm_par + solve_full) * 1e3
            sample = (f"C oracle on {cores} host cores (OpenMP threads standing in for MPI ranks): full FP64 "
                      f"assembly split over {cores} ranks ({round(asm_par * 1e3)} ms) + the whole FSILS_SOLVE "
                      f"(PRECONDDIAG and {sfull.itr} {what.split(' (')[0]}, suc {sfull.suc}) "
                      f"({round(solve_full * 1e3)} ms), timed in full")
        else:
            cpu_ms = (asm_par + (c2 - c1) + per_itc * itr64) * 1e3
            sample = (f"C oracle on {cores} host cores (OpenMP threads standing in for MPI ranks): full FP64 "
                      f"assembly split over {cores} ranks ({round(asm_par * 1e3)} ms) + PRECONDDIAG "
                      f"({round((c2 - c1) * 1e3)} ms) + {so.itr} {what} at {per_itc * 1e3:.2f} ms/iteration, "
                      f"scaled to the {itr64} iterations of the FP64 step")
        cpu1_ms = (asm1 + (c2 - c1) + per_it1 * itr64) * 1e3
        ref1 = None
        from oracle import refharness as rh

        if rh.available() and args.workload in ("fluid", "lelas"):
            # the reference's own Fortran (oracle/_ref), one MPI rank: its
            # element loop, and FSILS_SOLVE on a bounded sample of the same
            # system, scaled to the FP64 iteration count
            try:
                trac, bfl = prob.trac_faces, prob.bflow_faces
                prob.trac_faces, prob.bflow_faces = [], []
                _, _, _, _, ref_asm = rh.assemble(prob)
                prob.trac_faces, prob.bflow_faces = trac, bfl
                nodes = np.nonzero((mask == 0.0).any(axis=1))[0].astype(np.int32)
                RI = dict(relTol=smp["relTol"], absTol=smp["absTol"], mItr=smp["mItr"], sD=smp.get("sD", 0))
                _, rst, ref_solve = rh.solve(st.rowPtr - 1, st.colPtr - 1, dof, Vo, Ro, nodes, mask[nodes],
                                             smp["LS_type"], RI=RI)
                ref_it = ref_solve / max(rst["RI_itr"], 1)
                ref_ms = (ref_asm + ref_it * itr64) * 1e3
                ref1 = {"value": round(ref_ms, 1), "unit": "ms/step", "cores": 1, "kind": "reference",
                        "sample": (f"the reference's svFSI/svFSILS Fortran built from its sources (oracle/_ref), one "
                                   f"MPI rank: its element loop ({round(ref_asm * 1e3)} ms; face terms left out) + "
                                   f"FSILS_SOLVE for {rst['RI_itr']} iterations at {ref_it * 1e3:.2f} ms/iteration, "
                                   f"scaled to the {itr64} iterations of the FP64 step"),
                           "port_over_reference_1core": round(cpu1_ms / ref_ms, 3)}
                log(f"cpu leg: reference one rank {ref_ms:.0f} ms/step (port/reference {cpu1_ms / ref_ms:.2f})")
            except Exception as exc:  # the reference leg is informative only
                log(f"cpu leg: reference run skipped ({exc})")
        acc["assembly_rel_err_vs_cpu"] = float(np.abs(V0 - Vo).max() / np.abs(Vo).max())
        acc["residual_rel_err_vs_cpu"] = float(np.abs(R0 - Ro).max() / np.abs(Ro).max())
        for name, kern in struct_kernels if asm_kern else ():
            prob.quad = (base_quad & 1) | kern
            st.assemble()
            torch.cuda.synchronize()
            Vk = st.Val.cpu().numpy()
            asm_kern[name]["rel_err_vs_cpu"] = float(np.abs(Vk - Vo).max() / np.abs(Vo).max())
            del Vk
        prob.quad = base_quad
        log(f"cpu leg: solver samples in {c3 - c2:.1f} s; accuracy checks")
        # FP64 residual of each GPU solution, recomputed on the CPU: in the
        # reference's scaled norm |W (R - K x)| / |W R|, and unscaled,
        # |f - K u| / |f| over the dofs that are not Dirichlet-masked
        free = mask != 0
        for name, mixed in (("fp64", False), ("mixed", True)):
            st.assemble()
            st.solve(mixed=mixed)
            torch.cuda.synchronize()
            xg = st.R.cpu().numpy()
            y = np.where(W != 0, xg / np.where(W != 0, W, 1.0), 0.0)
            r = Rs - orc.spmv(st.rowPtr - 1, st.colPtr - 1, dof, Vs, y)
            acc[f"scaled_rel_residual_{name}"] = float(np.linalg.norm(r) / np.linalg.norm(Rs))
            ru = (Ro - orc.spmv(st.rowPtr - 1, st.colPtr - 1, dof, Vo, xg))[free]
            acc[f"unscaled_rel_residual_{name}"] = float(np.linalg.norm(ru) / np.linalg.norm(Ro[free]))
        cpu = {
            "value": round(cpu_ms, 1), "unit": "ms/step", "cores": cores, "kind": "port",
            "sample": sample,
            "single_core_value": round(cpu1_ms, 1),
            "single_core_sample": (f"one core: assembly {round(asm1 * 1e3)} ms, {so1.itr} iterations at "
                                   f"{per_it1 * 1e3:.2f} ms/iteration scaled to {itr64} (PRECONDDIAG as above, on "
                                   f"{cores} cores)"),
        }
        if ref1 is not None:
            cpu["reference_single_rank"] = ref1

    # assembly against HBM: Val written once, R, IEN, x and the state arrays
    # read once (zero-fill and face terms included in the timed call)
    k_state = 3 if args.workload == "struct" else 2
    asm_bytes = (nnzb * dof * dof * 8 + m.nNo * dof * 8 + 4 * m.nEl * m.eNoN + 24 * m.nNo
                 + k_state * m.nNo * st.Ag.shape[1] * 8)
    asm_gbs = asm_bytes / (asm_ms * 1e-3) / 1e9
    asm_roof = {"algorithmic_bytes": asm_bytes, "achieved_gbs": round(asm_gbs, 1),
                "frac_of_hbm_peak": round(asm_gbs / HBM_PEAK_GBS, 4),
                "note": "whole assembly call (zero-fill, element kernels, face terms) over the unique bytes"}
    speed = fp64_ms / mix_ms if (fp64_ms and mix_ms > 0) else None
    roof = None
    if gbs32 is not None:
        fused = dof == 4 and os.environ.get("SVFSI_FUSED_INNER", "0") == "1"
        poly = dof in (3, 4) and os.environ.get("SVFSI_INNER_POLY", "1") != "0"
        if amg_default:
            tt = 512 if nnzb >= (2 << 20) else 256
            fused = os.environ.get("SVFSI_AMG_FUSED", "") != "0"
            kname = (f"k_amg_resid_t<{dof}, {tt}> / k_amg_cheb_t<{dof}, {tt}" if fused else
                     f"k_sspmv<float, {dof}, {tt}")
            what = ("the finest level of the multigrid V-cycle: FP32 CSR-stream block SpMVs over row tiles "
                    + ("with the residual (r = b - A x) or the Chebyshev step fused in" if fused else
                       "of the Chebyshev smoothing and the residuals"))
        elif poly and os.environ.get("SVFSI_SPMV_STREAM", "") != "0":
            # the CSR-stream kernels over row tiles of 256 blocks (512 from 2M blocks on)
            tt = 512 if nnzb >= (2 << 20) else 256
            kname = f"k_spoly<float, {dof}, *, {tt}"
            what = ("polynomial preconditioner steps: FP32 CSR-stream block SpMV over row tiles with the "
                    "block-Jacobi and vector epilogues, modes 0/1/2 together, plus the setup Arnoldi SpMVs "
                    f"(k_sspmv<float, {dof}, {tt}>)")
        elif poly:
            nt = "true" if nnzb * dof * dof * 4 > (128 << 20) else "false"
            kname = f"k_poly{dof}<*, {nt}"
            what = ("polynomial preconditioner steps: FP32 block SpMV with the block-Jacobi and vector "
                    "epilogues, modes 0/1/2 together, plus the setup Arnoldi SpMVs")
        elif fused:
            nt = "true" if nnzb * 64 > (128 << 20) else "false"
            kname, what = f"k_gm_spmv_dots4<{nt}", "FP32 block SpMV fused with the inner multi-dot"
        else:
            kname = "k_bspmv4<float, 4" if dof == 4 else f"k_bspmv<{dof}, float, float"
            what = "FP32 block SpMV, inner GMRES"
        roof = {"kernel": f"{kname}> ({what})",
                "bound": "hbm", "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs32 / HBM_PEAK_GBS, 4), "traffic": measured_traffic(nnzb, kname),
                "bytes_per_launch": round(b32), "avg_launch_us": round(t32 / n32 * 1e3, 2), "launches": n32,
                "full_work_launches": n32_full, "timing": "events attached to each dispatch (hipExtLaunchKernelGGL)"}
        # the 256 MiB Infinity Cache (MALL) sits between L2 and HBM: a launch whose bytes fit it streams
        # partly from there, so its rate is an effective one and the FETCH/WRITE counters bound the HBM bytes
        # from above rather than measuring them
        mall = 256 << 20
        roof["working_set_bytes"] = round(b32)
        roof["residency"] = ("fits the 256 MiB Infinity Cache: rate against the HBM peak is an effective "
                             "rate, part of the bytes come from MALL" if b32 < mall else
                             "larger than the 256 MiB Infinity Cache: streams from HBM")
    out = {
        "metric": "Assembly+GMRES time per Newton step at 1e-8 residual; FP64 vs mixed; SpMV GB/s",
        "value": round(mix_ms, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mix_ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64+f32 (mixed: FP32 inner solve, FP64 residual/refinement)",
        "data": "synthetic",
        "config": {"workload": wl, "nNo": m.nNo, "nEl": m.nEl, "nnzb": nnzb, "dof": dof, "colors": st.ncolors},
        "fp64_ms_per_step": None if fp64_ms is None else round(fp64_ms, 3),
        "assembly_ms": round(asm_ms, 3),
        "assembly_roofline": asm_roof,
        **({"assembly_kernels": asm_kern} if asm_kern else {}),
        "speedup_mixed_vs_fp64": None if speed is None else round(speed, 3),
        **({} if poly_ms is None else {
            "inner_preconditioner": "aggregation multigrid (one FP32 V-cycle per outer step)",
            "mixed_polynomial_ms_per_step": round(poly_ms, 3),
            "mixed_polynomial_outer_itr": lsp[-1].RI.itr, "mixed_polynomial_spmv": lsp[-1].RI.cM,
            "speedup_polynomial_vs_nested_fp64": None if nest_ms is None else round(nest_ms / poly_ms, 3)}),
        "nested_fp64_ms_per_step": None if nest_ms is None else round(nest_ms, 3),
        "nested_fp64_outer_itr": None if lsn is None else lsn[-1].RI.itr,
        "nested_fp64_spmv": None if lsn is None else lsn[-1].RI.cM,
        "speedup_mixed_vs_nested_fp64": None if nest_ms is None else round(nest_ms / mix_ms, 3),
        "fp64_gmres_itr": None if ls64 is None else ls64[-1].RI.itr,
        "mixed_outer_itr": lsmx[-1].RI.itr,
        "mixed_inner_spmv": lsmx[-1].RI.cM,
        "suc": [None if ls64 is None else int(ls64[-1].RI.suc), int(lsmx[-1].RI.suc)],
        "spmv_fp64_gbs": None if gbs64 is None else round(gbs64, 1),
        "spmv_fp32_gbs": None if gbs32 is None else round(gbs32, 1),
        "accuracy": acc,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        out["speedup_vs_cpu_mixed"] = round(cpu["value"] / mix_ms, 1)
    print(json.dumps(out))
    return 0


def host_cores() -> int:
    """Host cores of this process's CPU share (one oracle thread per core,
    as the reference runs one MPI rank per core)."""
    cores = len(os.sched_getaffinity(0))
    return max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))


def bench_ns(args, prob, t0) -> int:
    """Assembly + FSILS_SOLVE with the NS solver (NSSOLVER.f, LS.f:69-95
    defaults: RI 0.4 / 10, GM 1e-2 / 2 restarts of 100, CG 0.2 / 500) per
    Newton step, FP64 on the GPU, beside the oracle's NSSOLVER on the host
    cores.  Roofline: the Schur-complement product (k_cgs_sp), the kernel
    that dominates the solve."""
    import ctypes as C

    import torch

    from svfsi_amd._lib import LS_TYPE_NS, lib
    from svfsi_amd.step import NewtonStep

    prob.ls = dict(LS_type=LS_TYPE_NS, relTol=0.4, absTol=1e-10, mItr=10)
    st = NewtonStep(prob)
    m = prob.mesh
    nnzb = st.nnz
    log(f"setup {time.time() - t0:.1f}s nNo={m.nNo} nEl={m.nEl} nnzb={nnzb} colors={st.ncolors}")
    for w in range(args.warmup):
        ls = st.step()
        log(f"warmup {w} ns: RI.itr {ls.RI.itr} GM.itr {ls.GM.itr} CG.itr {ls.CG.itr} suc {ls.RI.suc}")
    torch.cuda.synchronize()
    L = lib()
    L.svfsi_prof_enable(0)
    a = time.perf_counter()
    lss = [st.step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    ns_ms = (time.perf_counter() - a) * 1e3 / args.steps
    ls = lss[-1]
    a = time.perf_counter()
    for _ in range(args.steps):
        st.assemble()
    torch.cuda.synchronize()
    asm_ms = (time.perf_counter() - a) * 1e3 / args.steps
    log(f"ns step {ns_ms:.3f} ms (assembly {asm_ms:.3f} ms): RI.itr {ls.RI.itr} GM.itr {ls.GM.itr} "
        f"CG.itr {ls.CG.itr}")
    # profiled step: dispatch-attached events on the Schur product
    L.svfsi_prof_reset()
    L.svfsi_prof_enable(1)
    lsp = st.step()
    torch.cuda.synchronize()
    L.svfsi_prof_enable(0)
    cnt, tot, byt = C.c_longlong(), C.c_double(), C.c_double()
    L.svfsi_prof_get(4, C.byref(cnt), C.byref(tot))
    L.svfsi_prof_get_bytes(4, C.byref(byt))
    full = max(1, min(cnt.value, lsp.CG.itr))  # launches queued after convergence exit at once
    bpl = byt.value / max(cnt.value, 1)
    gbs = bpl * full / (tot.value * 1e-3) / 1e9 if cnt.value else None
    # the row-tile kernel from 2M blocks on (512-block tiles), the 8-lanes-per-row one below
    ct = os.environ.get("SVFSI_CGS_TILE")
    ct = int(ct) if ct in ("0", "512", "1024") else (512 if nnzb >= (2 << 20) else 0)
    kname = f"k_cgs_sp_t<{ct}>" if ct else "k_cgs_sp<8>"
    roof = None if gbs is None else {
        "kernel": f"{kname} (Schur-complement product SP = L P - Gt GP with its partial dot)", "bound": "hbm",
        "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(nnzb, kname), "bytes_per_launch": round(bpl),
        "avg_launch_us": round(tot.value / full * 1e3, 2), "launches": cnt.value, "full_work_launches": full,
        "timing": "events attached to each dispatch (hipExtLaunchKernelGGL)"}
    cpu = None
    acc = {}
    if not args.no_cpu:
        from oracle import oracle as orc

        st.assemble()
        torch.cuda.synchronize()
        V0 = st.Val.cpu().numpy().copy()
        R0 = st.R.cpu().numpy().copy()
        st.solve()
        torch.cuda.synchronize()
        xg = st.R.cpu().numpy()
        cores = host_cores()
        rp, cp = st.rowPtr - 1, st.colPtr - 1
        log("cpu leg: oracle assembly and NSSOLVER")
        c0 = time.perf_counter()
        Vo, Ro = orc.assemble(prob, rp, cp)
        asm1 = time.perf_counter() - c0
        asm_par = orc.assemble_ranks_seconds(prob, rp, cp, cores)
        t_solve = {}
        for nt in (1, cores):
            orc.set_threads(nt)
            c0 = time.perf_counter()
            xo, _, lo = orc.fsils_solve_ns(rp, cp, 4, prob.dir_mask(), prob.ls, R0, V0)
            t_solve[nt] = time.perf_counter() - c0
        orc.set_threads(1)
        acc["assembly_rel_err_vs_cpu"] = float(np.abs(V0 - Vo).max() / np.abs(Vo).max())
        acc["solution_rel_diff_vs_cpu"] = float(np.abs(xg - xo).max() / np.abs(xo).max())
        acc["cpu_itr"] = [lo.RI.itr, lo.GM.itr, lo.CG.itr]
        Vs, Rs, W = orc.precond_diag(rp, cp, 4, V0, R0, prob.dir_mask())
        for name, x in (("gpu", xg), ("cpu", xo)):
            y = np.where(W != 0, x / np.where(W != 0, W, 1.0), 0.0)
            r = Rs - orc.spmv(rp, cp, 4, Vs, y)
            acc[f"scaled_rel_residual_{name}"] = float(np.linalg.norm(r) / np.linalg.norm(Rs))
        cpu_ms = (asm_par + t_solve[cores]) * 1e3
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/step", "cores": cores, "kind": "port",
               "sample": (f"C oracle on {cores} host cores (OpenMP threads standing in for MPI ranks): full FP64 "
                          f"assembly split over {cores} ranks ({round(asm_par * 1e3)} ms) + the whole NSSOLVER "
                          f"solve with PRECONDDIAG ({round(t_solve[cores] * 1e3)} ms)"),
               "single_core_value": round((asm1 + t_solve[1]) * 1e3, 1),
               "single_core_sample": (f"the same on one core: assembly {round(asm1 * 1e3)} ms, solve "
                                      f"{round(t_solve[1] * 1e3)} ms")}
    wl = ("FLUID3D VMS tet4 cylinder r=1 L=10 (h=%g), one Newton step, NSSOLVER with the reference defaults "
          "(RI 0.4/10, GM 1e-2/2x100, CG 0.2/500) and diagonal scaling" % (0.042 if args.workload == "fluid-large"
                                                                            else args.h))
    out = {
        "metric": "Assembly + NS-solver time per Newton step (FP64)",
        "value": round(ns_ms, 3), "unit": "ms/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ns_ms, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl, "nNo": m.nNo, "nEl": m.nEl, "nnzb": nnzb, "dof": 4, "colors": st.ncolors},
        "assembly_ms": round(asm_ms, 3),
        "ns_itr": {"RI": ls.RI.itr, "GM": ls.GM.itr, "CG": ls.CG.itr, "suc": int(ls.RI.suc),
                   "fNorm_over_iNorm": ls.RI.fNorm / ls.RI.iNorm, "Resm": ls.Resm, "Resc": ls.Resc},
        "accuracy": acc, "roofline": roof, "cpu_baseline": cpu,
    }
    if cpu is not None:
        out["speedup_vs_cpu"] = round(cpu["value"] / ns_ms, 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
