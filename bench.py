"""Benchmark: one 2CLOWNS-3D pressure solve per step on one MI355X.

A step is one pass of the hot path over the configuration-2 workload
(256x256x64 cells, wavy free surface at 0.6 H, 64x64x24 solid obstacle,
seeded U(-0.1,0.1) face velocities): device assembly of the pressure system
fused with PCGPME's row scaling and sort (init_matrix + cal_matrix_element +
RHS + pcgpme_para.F90:77-112), then the PM-preconditioned Bi-CGSTAB solve to
the reference's stopping rule sqrt(RN/BN) <= 1e-10, FP64 end to end.  The
geometry (classified once on the host) is resident in HBM before timing; no
state carries from one step to the next (x0 = 0, operator rebuilt).

The same steps are then run in mixed precision (FP32 inner solve, FP64
refinement), reported under "mixed" with the error against the FP64 solution.

Prints one JSON line.  `value` is seconds per step (lower is better).

`--will NXxNY` measures the second call site of the same solver instead: the
2DH dispersion solve PCGPME_WILL (src2LOWNS/pcgpme_2DH.f90, called once per
time step by disp_adjust, 1sub-2DH-momentum.f90:325-326) on the 2DH
linear-dispersion matrix of an NX x NY grid (deep water, 400-500 m cells).
The matrix is constant in time (disp_param builds it once), so it stays
resident; a step is one solve_will on a new right-hand side already in HBM
(solve + the reference's output pass).  The drop-in (host arrays in and out,
matrix upload included) is timed beside it.

`--vof 2|3` measures the VOF update cal_vof (src2LOWNS/1sub-vof.F90:10-38,
up to its closing set_nflag) on configuration 2 or 3: face fluxes with the
quvwf_kaidan surface stage, quvwf_modify, cal_new_f, simple_slide(0).  The
classified grid, the face velocities and F of the previous step are
resident; a step is one update (clowns_solver_vof).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2clowns-3d_amd"))

PEAK_HBM_GBS = 8000.0
DESCR = {1: "wavy surface at 0.6H + 16x16x12 obstacle", 2: "wavy surface at 0.6H + 64x64x24 obstacle",
         3: "40 Gaussian bed bumps, partial cells fb in [0.05,1], wavy surface at 0.6H"}
# Algorithmic bytes per matrix row (values and vectors each kernel must stream
# once; index arrays excluded).  Whole iteration: 512 B/row at FP64.
ALG_BYTES = {
    "spmv": 72,        # AL 48 + D 8 + x 8 + y 8
    "pm_stage": 24,    # AL_j 8 + in 8 + out 8
    "pm_pair": 32,     # fused pair of stages: AL_j 8 + AL_j+1 8 + in 8 + out 8
    "iteration": 512,
    "setup_cold": 288,
    "iteration32": 256,   # FP32 inner iteration of the mixed solve
    "refine64": 216,      # FP64 true residual: SpMV with b - Ax (80) + one PM application (136)
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3])
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=40)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only assemble and run the timed kernel loops (for rocprofv3 --pmc)")
    ap.add_argument("--will", default=None, metavar="NXxNY",
                    help="measure the 2DH dispersion solve PCGPME_WILL on an NX x NY grid instead")
    ap.add_argument("--vof", type=int, default=None, choices=[2, 3],
                    help="measure the VOF update cal_vof on configuration 2 or 3 instead")
    return ap.parse_args()


# Algorithmic bytes per row of one PCGPME_WILL iteration (NL = 4): 2 SpMV
# (AL 32 + D 8 + x 8 + y 8 = 56, the e-fused one 72), 8 single-term PM stages
# (24 each) plus the two dot vectors read by the last stage of each (16),
# p update (32), x/r update with the dots (56).
WILL_ITER_BYTES = 56 + 72 + 8 * 24 + 16 + 32 + 56


def will_cpu_baseline(s, threads, eps, iters):
    """The reference's own PCGPME_WILL (flang build in oracle/_ref) on the
    same system at `threads` OpenMP threads; the C oracle if that was not built."""
    import numpy as np
    N, NA = s["N"], s["NA"]
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_pcgpme.so")
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    AL, LL, D = s["AL"], s["LL"], s["D"]
    times = []
    for _ in range(2):
        B = s["B"].copy()
        lib = None
        if os.path.exists(ref):
            lib = ctypes.CDLL(ref)
            kind = "reference"
            i32 = lambda v: ctypes.byref(ctypes.c_int(v))
            f64 = lambda v: ctypes.byref(ctypes.c_double(v))

            def call():
                lib.ref_set_threads(threads)
                lib.pcgpme_will_(P(AL), i32(4), i32(NA), P(LL), P(D), i32(N), i32(64), P(B), f64(eps), f64(5.0),
                                 i32(iters))
        else:
            orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
            kind = "port"
            ier, k = ctypes.c_int(), ctypes.c_int()

            def call():
                orc.oracle_set_threads(threads)
                orc.oracle_pcgpme_will(P(AL), 4, NA, P(LL), P(D), N, 64, P(B), ctypes.c_double(eps),
                                       ctypes.c_double(5.0), iters, ctypes.byref(ier), ctypes.byref(k), None, None)
        out = {}

        def body():
            t0 = time.perf_counter()
            call()
            out["t"] = time.perf_counter() - t0

        old = threading.stack_size()
        threading.stack_size(max(32 << 20, 200 * (N + 64) + 64 * NA + (64 << 20)))
        th = threading.Thread(target=body)
        with _quiet_stdout(lib):
            th.start()
            th.join()
        threading.stack_size(old)
        times.append(out["t"])
    return {"value": min(times), "unit": "s/solve", "cores": threads, "kind": kind,
            "sample": (f"the same 2DH system (N={N}) on the host; "
                       f"{'flang build of pcgpme_2DH.f90' if kind == 'reference' else 'C oracle of PCGPME_WILL'} "
                       f"run to convergence at {threads} OpenMP threads, best of 2")}


def bench_will(args):
    import numpy as np
    import torch
    import clowns3d
    from clowns3d import Solver, configs
    nx, ny = (int(v) for v in args.will.lower().split("x"))
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    s = configs.dispersion_2dh(nx, ny, depth=4000.0, dx=(400.0, 500.0))
    N = s["N"]
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol = Solver(0)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    if args.profile_kernels:
        sol.solve_will(t(s["B"]), eps=args.eps, iters=1000)
        ix = "Ix%d" % sol.index_bits()
        st = sol.tuning()["stream64"]
        sp, sm = ("true" if st > 0 and st & 2 else "false"), ("true" if st > 0 and st & 1 else "false")
        res = {k: sol.time_kernel(i, 10) * 1e3 for k, i in
               ((f"k_spmv_v2<double,4,0,{ix},{sp}>", 0), (f"k_pm_v2<double,1,0,{ix},{sm},double>", 1), ("k_copy<double>", 4))}
        print(json.dumps({"rows": N, "avg_us": res}))
        return
    rng = np.random.default_rng(7)
    # a new right-hand side per step: the base PSI scaled, as the surface evolves
    rhs = [t(s["B"] * rng.uniform(0.5, 1.5)) for _ in range(args.steps + args.warmup)]
    out = torch.empty(N, dtype=torch.float64, device=dev)
    for k in range(args.warmup):
        sol.solve_will(rhs[k], out, eps=args.eps, iters=1000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [sol.solve_will(rhs[args.warmup + k], out, eps=args.eps, iters=1000)[1] for k in range(args.steps)]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    assert all(i["ier"] == 0 for i in infos), infos
    its = infos[-1]["iters"]
    # drop-in: host arrays, matrix upload, copy back; the second of two calls
    # (the first includes the library solver's one-time form tuning)
    for _ in range(2):
        c = {k: (v.copy(order="F") if isinstance(v, np.ndarray) else v) for k, v in s.items()}
        t1 = time.perf_counter()
        ier = clowns3d.pcgpme_will(c["AL"], 4, c["NA"], c["LL"], c["D"], N, 64, c["B"], args.eps, 5.0, 1000)
        t_drop = time.perf_counter() - t1
    # dominant kernel, HIP events on the solver's stream
    # per iteration: 2 SpMV and two NL = 4 PM applications, as 8 single
    # stages or (fused) 2 pairs + 4 single stages, whichever the solver picked
    t_spmv = sol.time_kernel(0, 50)
    t_pm = sol.time_kernel(1, 50)
    t_pair = sol.time_kernel(5, 50)
    t_copy = sol.time_kernel(4, 50)
    ix = "Ix%d" % sol.index_bits()
    tune = sol.tuning()
    st = tune["stream64"]
    sp, sm = ("true" if st > 0 and st & 2 else "false"), ("true" if st > 0 and st & 1 else "false")
    cands = [(2 * t_spmv, f"k_spmv_v2<double,4,0,{ix},{sp}>", t_spmv, 56 * N)]
    if tune["fuse64"] == 1:
        cands.append((2 * t_pair, f"k_pm2_v2<double,{ix},{sm},double>", t_pair, ALG_BYTES["pm_pair"] * N))
        cands.append((4 * t_pm, f"k_pm_v2<double,1,0,{ix},{sm},double>", t_pm, ALG_BYTES["pm_stage"] * N))
    else:
        cands.append((8 * t_pm, f"k_pm_v2<double,1,0,{ix},{sm},double>", t_pm, ALG_BYTES["pm_stage"] * N))
    _, kname, kms, kbytes = max(cands)
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"traffic_will{nx}x{ny}.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get(kname)
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = will_cpu_baseline(s, threads, args.eps, 1000)
    solve_gbs = N * its * WILL_ITER_BYTES / (infos[-1]["ms"] * 1e-3) / 1e9
    res = {
        "metric": "2DH dispersion solve (PCGPME_WILL) s/solve to 1e-10 rel. residual; HBM GB/s",
        "value": dt, "unit": "s/solve", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2DH linear-dispersion matrix {nx}x{ny} (depth 4000 m, 400-500 m cells, "
                               f"Gaussian surface hump), resident operator, PCGPME_WILL to {args.eps:g} + output pass",
                   "rows": N, "iterations": its, "filter_rounds": infos[-1]["filter_rounds"],
                   "solve_ms_hip_events": infos[-1]["ms"], "solve_alg_GBps": solve_gbs,
                   "solve_alg_frac": solve_gbs / PEAK_HBM_GBS,
                   "drop_in_ms_host_arrays": t_drop * 1e3, "drop_in_ier": ier},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kname,
                     "alg_bytes_per_launch": kbytes, "avg_us": kms * 1e3,
                     "copy_GBps": 16 * N / (t_copy * 1e-3) / 1e9},
        "cpu_baseline": cpu,
    }
    if cpu:
        res["speedup_vs_cpu"] = cpu["value"] / dt
    print(json.dumps(res))


# Algorithmic bytes per box cell of the two streaming VOF kernels (arrays read
# once, written once; neighbour values come from cache).  Open cells:
#   face fluxes  un 24 + F 8 + nf 1 + nb 1 + a_x,a_y 16 (kaidan) + qf 24 = 74
#   new F        qf 24 + a 24 + fb 8 + Fp 8 + nf 1 + F_new 8 = 73
# solid cells: flux nf 1 + qf 24 = 25; new F nf 1 + F 8 + F_new 8 = 17.
VOF_BYTES = {"flux": (74, 25), "newf": (73, 17)}


def vof_cpu_baseline(g, A, Fp, dt, threads):
    """The reference's own VOF routines (flang build in oracle/_ref, after its
    set_nflag at one thread, untimed) at `threads` OpenMP threads; the C
    oracle (serial) if that was not built."""
    import numpy as np
    nc = g.ncell
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_vof.so")
    times = []
    for _ in range(2):
        qf, fo, fc = np.zeros(3 * nc), np.zeros(nc), np.zeros(nc)
        nf, nb = np.zeros(nc, np.int8), np.zeros(nc, np.int8)
        dtv, secs = ctypes.c_double(dt), ctypes.c_double(0.0)
        if os.path.exists(ref):
            lib = ctypes.CDLL(ref, mode=os.RTLD_LAZY)
            kind = "reference"
            qs, mask = np.zeros(3 * nc), np.zeros(3 * nc, np.int8)
            with _quiet_stdout():
                lib.ref_vof(g.nx, g.ny, g.nz, P(A["solid"]), P(A["F"]), P(Fp), P(A["fb"]), P(A["a"]), P(A["un"]),
                            P(A["dx"]), P(A["xf"]), g.stride, ctypes.byref(dtv), 1, P(qs), P(mask), P(nf), P(nb),
                            P(fc), P(qf), P(fo), threads, ctypes.byref(secs))
            times.append(secs.value)
        else:
            from clowns3d import assembly
            c = assembly.classify(g)
            orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
            kind, threads = "port", 1
            retries = ctypes.c_int(0)
            t0 = time.perf_counter()
            orc.oracle_vof(g.nx, g.ny, g.nz, g.stride, P(A["solid"]), P(c["nf"]), P(c["nb"]), P(c["F"]), P(Fp),
                           P(A["fb"]), P(A["a"]), P(A["un"]), P(A["dx"]), ctypes.byref(dtv), 1, None, None, 0, None,
                           None, P(qf), P(fo), ctypes.byref(retries), None)
            times.append(time.perf_counter() - t0)
    return {"value": min(times), "unit": "s/step", "cores": threads, "kind": kind,
            "sample": (f"the same grid ({nc} cells) on the host; "
                       f"{'flang build of cal_quvwf, quvwf_kaidan, quvwf_modify, cal_new_f, simple_slide' if kind == 'reference' else 'C oracle of cal_vof'}"
                       f" timed around the update at {threads} OpenMP threads, best of 2")}


def bench_vof(args):
    import numpy as np
    import torch
    from clowns3d import DeviceGrid, Solver, assembly, configs
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = configs.CONFIGS[args.vof]()
    c = assembly.classify(g)
    A = c["arrays"]
    dg = DeviceGrid(g, classified=c)
    sol = Solver(0)
    Fp_h = c["F"].copy()
    Fp = torch.from_numpy(Fp_h).to(dev)
    out = torch.empty(g.ncell, dtype=torch.float64, device=dev)
    nc = g.ncell
    n_open = int((A["solid"] == 0).sum())
    if args.profile_kernels:
        dg.vof(sol, Fp, mode=1, out_F=out)
        res = {k: dg.vof_time_kernel(sol, Fp, i, 10) * 1e3 for k, i in
               (("k_vof_flux<1>", 0), ("k_vof_newf", 1), ("k_vof_slide", 2), ("k_copy<double>", 3))}
        print(json.dumps({"cells": nc, "avg_us": res}))
        return
    for _ in range(args.warmup):
        dg.vof(sol, Fp, mode=1, out_F=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = [dg.vof(sol, Fp, mode=1, out_F=out)[2] for _ in range(args.steps)]
    torch.cuda.synchronize()
    dt_step = (time.perf_counter() - t0) / args.steps
    info = infos[-1]
    kt = {k: dg.vof_time_kernel(sol, Fp, i, 50) for k, i in (("flux", 0), ("newf", 1), ("slide", 2), ("copy", 3))}
    kname, key = ("k_vof_flux<1>", "flux") if kt["flux"] >= kt["newf"] else ("k_vof_newf", "newf")
    bo, bs = VOF_BYTES[key]
    kbytes = n_open * bo + (nc - n_open) * bs
    achieved = kbytes / (kt[key] * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"traffic_vof_config{args.vof}.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get(kname)
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = vof_cpu_baseline(g, A, Fp_h, g.dt, threads)
    res = {
        "metric": "VOF update (cal_vof) s/step; HBM GB/s",
        "value": dt_step, "unit": "s/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt_step * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"configuration {args.vof}: {g.name}, {DESCR[args.vof]}; VOF update with the "
                               "quvwf_kaidan surface stage, classified grid resident",
                   "cells": nc, "open_cells": n_open, "modified": info["modified"], "slid": info["slid"],
                   "rounds": info["rounds"], "retries": info["retries"], "update_ms_hip_events": info["ms"],
                   "kernel_us": {k: v * 1e3 for k, v in kt.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kname,
                     "alg_bytes_per_launch": kbytes, "avg_us": kt[key] * 1e3,
                     "copy_GBps": 16 * nc / (kt["copy"] * 1e-3) / 1e9},
        "cpu_baseline": cpu,
    }
    if cpu:
        res["speedup_vs_cpu"] = cpu["value"] / dt_step
    print(json.dumps(res))


class _quiet_stdout:
    """Route file descriptor 1 to /dev/null around a call into the flang
    build of the reference, which writes progress lines to unit 6 (flushed
    before the descriptor is restored), so stdout carries only the JSON line."""

    def __init__(self, lib=None):
        self.lib = lib

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)
        return self

    def __exit__(self, *exc):
        if self.lib is not None and hasattr(self.lib, "ref_flush"):
            self.lib.ref_flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)
        return False


def cpu_baseline(g, iters_full, n_iter_hi, threads):
    """The reference's own PCGPME (flang build in oracle/_ref) -- or, if that
    was not built, the C oracle -- on the same configuration's system built
    on the host, run for 5 and n_iter_hi iterations; solve time =
    setup + iters_full * per-iteration time."""
    import numpy as np
    orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    A = g.arrays()
    nc = g.ncell
    nf = np.zeros(nc, np.int8)
    nb = np.zeros(nc, np.int8)
    r2c = np.zeros(nc, np.int32)
    mno = np.zeros(nc, np.int32)
    lda = nc
    AL = np.zeros((6, lda))
    LL = np.zeros((6, lda), np.int32)
    D = np.zeros(lda)
    B = np.zeros(lda)
    orc.oracle_build_system.restype = ctypes.c_int
    N = orc.oracle_build_system(g.nx, g.ny, g.nz, P(A["solid"]), P(A["F"]), P(A["fb"]), P(A["a"]), P(A["un"]),
                                P(A["dx"]), P(A["xf"]), g.stride, ctypes.c_double(g.dt), P(nf), P(nb), P(r2c),
                                P(mno), lda, P(AL), P(LL), P(D), P(B))
    NA = N + 60
    ALf = np.asfortranarray(AL[:, :NA].T)
    LLf = np.asfortranarray(LL[:, :NA].T)
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_pcgpme.so")
    kind = "reference" if os.path.exists(ref_path) else "port"

    def run(iters):
        a, l, d, b = ALf.copy(order="F"), LLf.copy(order="F"), D[:N].copy(), B[:N].copy()
        x = np.zeros(N + 64)
        eps = np.array([1e-10, 0.0])
        it = ctypes.c_int(iters)
        ier = ctypes.c_int(0)
        lib = None
        if kind == "reference":
            lib = ctypes.CDLL(ref_path)
            i32 = lambda v: ctypes.byref(ctypes.c_int(v))

            def call():
                # the OpenMP thread count is a per-thread setting: set it in
                # the thread that calls the reference
                lib.ref_set_threads(threads)
                lib.pcgpme_(P(a), i32(6), i32(NA), P(l), P(d), i32(N), i32(64), P(b), P(x), P(eps),
                            ctypes.byref(ctypes.c_double(5.0)), ctypes.byref(it), ctypes.byref(ier))
        else:
            orc.oracle_set_threads(threads)
            call = lambda: orc.oracle_pcgpme(P(a), 6, NA, P(l), P(d), N, 64, P(b), P(x), P(eps),
                                             ctypes.c_double(5.0), ctypes.byref(it), ctypes.byref(ier))
        out = {}

        def body():
            t0 = time.perf_counter()
            call()
            out["t"] = time.perf_counter() - t0

        old = threading.stack_size()
        # the reference's work arrays are automatic (stack) arrays
        threading.stack_size(max(32 << 20, 128 * (N + 64) + 32 * NA + (64 << 20)))
        th = threading.Thread(target=body)
        with _quiet_stdout(lib):
            th.start()
            th.join()
        threading.stack_size(old)
        return out["t"], it.value

    lo = 5
    t_lo, _ = run(lo)
    t_hi, _ = run(n_iter_hi)
    t_it = (t_hi - t_lo) / (n_iter_hi - lo)
    t_setup = max(0.0, t_lo - lo * t_it)
    value = t_setup + iters_full * t_it
    return {"value": value, "unit": "s/step", "cores": threads, "kind": kind,
            "sample": (f"configuration {g.name} system (N={N}) built on the host; "
                       f"{'flang build of pcgpme_para.F90' if kind == 'reference' else 'C oracle of PCGPME'} "
                       f"timed for {lo} and {n_iter_hi} iterations at {threads} OpenMP threads "
                       f"({t_it * 1e3:.1f} ms/iteration, setup {t_setup * 1e3:.0f} ms); "
                       f"value = setup + {iters_full} iterations (the GPU FP64 count)"),
            "ms_per_iteration": t_it * 1e3}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the rows into z-slabs, one per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.will:
        bench_will(args)
        return
    if args.vof:
        bench_vof(args)
        return
    import numpy as np
    import torch
    import clowns3d
    from clowns3d import DeviceGrid, Solver, configs

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = configs.CONFIGS[args.config]()
    dg = DeviceGrid(g)
    N = dg.N
    sol = Solver(0)
    b = torch.empty(N, dtype=torch.float64, device=dev)
    x = torch.zeros(N, dtype=torch.float64, device=dev)

    if args.profile_kernels:
        # the kernels in the forms the solver picks for this operator
        dg.assemble_into(sol, b=b)
        sol.solve(b, x, eps=args.eps)
        sol.solve(b, x, eps=args.eps, mixed=True)
        tune = sol.tuning()
        ix, ix32 = "Ix%d" % tune["index64"], "Ix%d" % tune["index32"]
        # the FP32 inner preconditioner's coefficient type (CLOWNS_PM_HALF)
        c32 = "float" if os.environ.get("CLOWNS_PM_HALF", "1") == "0" else "_Float16"
        v4 = os.environ.get("CLOWNS_V4_32", "1") != "0"  # FP32 single stages: four rows a thread
        rows = ix32 == "Ix16" and os.environ.get("CLOWNS_SPMV_ROWS", "1") != "0"  # FP32 SpMV, row-major
        st = {p: ("true" if tune["stream" + p] > 0 and tune["stream" + p] & 1 else "false",
                  "true" if tune["stream" + p] > 0 and tune["stream" + p] & 2 else "false") for p in ("64", "32")}
        res = {k: sol.time_kernel(i, 10) * 1e3 for k, i in
               ((f"k_spmv_v2<double,6,0,{ix},{st['64'][1]}>", 0), (f"k_pm_v2<double,1,0,{ix},{st['64'][0]},double>", 1),
                ("k_copy<double>", 4), (f"k_pm2_v2<double,{ix},{st['64'][0]},double>", 5),
                (f"k_spmv_rows32<{st['32'][1]}>" if rows else f"k_spmv_v2<float,6,0,{ix32},{st['32'][1]}>", 2), (f"k_pm_v4<{ix32},{st['32'][0]},{c32}>" if v4 else f"k_pm_v2<float,1,0,{ix32},{st['32'][0]},{c32}>", 3),
                (f"k_pm2_v2<float,{ix32},{st['32'][0]},{c32}>", 6))}
        print(json.dumps({"rows": N, "avg_us": res, "tuning": tune}))
        return

    def step(mixed):
        dg.assemble_into(sol, b=b)
        x.zero_()
        _, info = sol.solve(b, x, eps=args.eps, maxiter=10000, mixed=mixed)
        return info

    def run(mixed):
        for _ in range(args.warmup):
            step(mixed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [step(mixed) for _ in range(args.steps)]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return (t1 - t0) / args.steps, infos

    t64, infos64 = run(False)
    x64 = x.clone()
    it64 = infos64[-1]["iters"]
    assert all(i["ier"] == 0 for i in infos64), infos64

    # accuracy reference for both forms: a tight FP64 solve (stagnation exit)
    dg.assemble_into(sol, b=b)
    xt = torch.zeros_like(x)
    sol.solve(b, xt, eps=2e-15, maxiter=20000)
    err64 = ((x64 - xt).norm() / xt.norm()).item()
    mixed = None
    if not args.no_mixed:
        t32, infos32 = run(True)
        i32 = infos32[-1]
        mixed_bytes = N * (i32["iters"] * ALG_BYTES["iteration32"] + (i32["refinements"] + 1) * ALG_BYTES["refine64"])
        mixed_gbs = mixed_bytes / (i32["ms"] * 1e-3) / 1e9
        mixed = {"value": t32, "ms_per_step": t32 * 1e3, "inner_iters": i32["iters"],
                 "refinements": i32["refinements"], "err": i32["err"], "ier": i32["ier"],
                 "solve_ms_hip_events": i32["ms"], "solve_alg_GBps": mixed_gbs,
                 "solve_alg_frac": mixed_gbs / PEAK_HBM_GBS,
                 "rel_l2_vs_tight_fp64": ((x - xt).norm() / xt.norm()).item(),
                 "rel_l2_vs_fp64": ((x - x64).norm() / x64.norm()).item(), "speedup_vs_fp64": t64 / t32}

    # dominant kernel of the FP64 iteration, timed live with HIP events on
    # the solver's stream.  Per iteration: 2 SpMV, and 8 single-term PM stages
    # or 4 fused pairs (whichever the solver times faster on this operator,
    # as it does itself).  A fused pair's algorithmic bytes are its own
    # minimum (two coefficient columns, the input, the output: 32 B/row), not
    # the two passes it replaces.
    dg.assemble_into(sol, b=b)
    sol.solve(b, x, eps=args.eps)
    t_spmv = sol.time_kernel(0, 50)
    t_pm = sol.time_kernel(1, 50)
    t_pair = sol.time_kernel(5, 50)
    t_copy = sol.time_kernel(4, 50)
    ix = "Ix%d" % sol.index_bits()
    tune = sol.tuning()
    st_pm = "true" if tune["stream64"] > 0 and tune["stream64"] & 1 else "false"
    st_sp = "true" if tune["stream64"] > 0 and tune["stream64"] & 2 else "false"
    cands = [(2 * t_spmv, f"k_spmv_v2<double,6,0,{ix},{st_sp}>", t_spmv, ALG_BYTES["spmv"] * N)]
    if tune["fuse64"] == 1:
        cands.append((4 * t_pair, f"k_pm2_v2<double,{ix},{st_pm},double>", t_pair, ALG_BYTES["pm_pair"] * N))
    else:
        cands.append((8 * t_pm, f"k_pm_v2<double,1,0,{ix},{st_pm},double>", t_pm, ALG_BYTES["pm_stage"] * N))
    _, kname, kms, kbytes = max(cands)
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"traffic_config{args.config}.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get(kname)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kname,
                "alg_bytes_per_launch": kbytes, "avg_us": kms * 1e3,
                "copy_GBps": 16 * N / (t_copy * 1e-3) / 1e9}
    solve_ms = infos64[-1]["ms"]
    solve_gbs = N * (ALG_BYTES["setup_cold"] + it64 * ALG_BYTES["iteration"]) / (solve_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(g, it64, args.cpu_iters, threads)

    out = {
        "metric": "pressure-solve s/step to 1e-10 rel. residual, FP64 vs mixed-prec; HBM GB/s",
        "value": t64, "unit": "s/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t64 * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"configuration {args.config}: {g.name} ({DESCR[args.config]}), "
                               f"assemble + PM-Bi-CGSTAB to sqrt(RN/BN) <= {args.eps:g}",
                   "rows": N, "cells": g.nx * g.ny * g.nz, "iterations": it64,
                   "err": infos64[-1]["err"], "rel_l2_vs_tight_fp64": err64, "solve_ms_hip_events": solve_ms,
                   "solve_alg_GBps": solve_gbs, "solve_alg_frac": solve_gbs / PEAK_HBM_GBS},
        "mixed": mixed,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = cpu["value"] / t64
    print(json.dumps(out))


if __name__ == "__main__":
    main()
