"""Benchmark: one 2CLOWNS-3D pressure solve per step on one MI355X.

A step is one pass of the hot path over the configuration-3 workload (by
default; --config 1|2 picks the smaller ones): 512x512x96 cells, 40 seeded
Gaussian bed bumps with partial cells (fb in [0.05, 1]), wavy free surface at
0.6 H, seeded U(-0.1,0.1) face velocities -- 13.07M matrix rows.  The step is
the device assembly of the pressure system fused with PCGPME's row scaling and
sort (init_matrix + cal_matrix_element + RHS + pcgpme_para.F90:77-112), then
the PM-preconditioned Bi-CGSTAB solve to the reference's stopping rule
sqrt(RN/BN) <= 1e-10, FP64 end to end (`value`).  The geometry (classified
once on the host) is resident in HBM before timing; no state carries from one
step to the next (x0 = 0, operator rebuilt).

The same steps then run in mixed precision (FP32 inner solve, FP64
refinement), reported under "mixed" with its error against a tight FP64
solve and against the FP64 solution, and its own roofline entry.

roofline: one more step of each form with HIP events around every kernel
launch on the solver's stream (clowns_solver_kernel_profile): the kernel
class with the most in-solve time, its average launch duration and its
algorithmic bytes per launch (CLASS_BYTES x rows); `traffic` from
profiles/traffic_config<n>.json, written by --traffic from two rocprofv3
--pmc runs (FETCH_SIZE, WRITE_SIZE) of --pmc-run.

cpu_baseline: the reference's own PCGPME (flang build of pcgpme_para.F90,
oracle/_ref) on the same system assembled on the host, run to convergence at
the host's cores (OMP_PROC_BIND=close), median of --cpu-runs after a warm-up,
with the reference's ITER and the relative L2 distance of each GPU solution
from the reference's X.  boundary: the drop-in clowns_pcgpme with the
caller's host arrays (PCIe in and out) beside the resident solve.

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
import itertools
import json
import math
import os
import subprocess
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
    "refine64": 240,      # per FP64 refinement of the mixed solve (true residual, x update, downcast)
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=3, choices=[1, 2, 3])
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-boundary", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--pmc-run", default=None, metavar="DIR",
                    help="solves in every kernel form plus the copy kernel, for rocprofv3 --pmc (row count to DIR); "
                         "with --will, PCGPME_WILL solves")
    ap.add_argument("--traffic", nargs=2, default=None, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write profiles/traffic_config<n>.json (traffic_will<NXxNY>.json with --will) from two "
                         "rocprofv3 --pmc runs of --pmc-run")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="--vof: only run the timed kernel loops (for rocprofv3 --pmc)")
    ap.add_argument("--reference-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--trace-summary", default=None, metavar="CSV",
                    help="print a markdown summary of a rocprofv3 --kernel-trace CSV (working launches)")
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
# per-launch algorithmic bytes per row of PCGPME_WILL's kernel classes (NL = 4)
WILL_CLASS_BYTES = {"spmv": 56, "pm_stage": 24, "pm_pair": 32, "pm_dot": 32, "p_update": 32, "e_update": 24,
                    "xr_update": 56}


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
    if args.pmc_run:
        # solves in each kernel form the solver may pick for this operator
        # (streamed or cached arrays, 8- or 16-bit codes, single or fused
        # PM stages), then the copy kernel the counters are scaled by
        tunes = []
        for st, ll8, fuse in itertools.product(("0", "1", "2", "3"), ("0", "1"), ("0", "1")):
            env = {"CLOWNS_STREAM": st, "CLOWNS_LL8": ll8, "CLOWNS_FUSE_PM": fuse}
            os.environ.update(env)
            try:
                s2 = Solver(0)
            finally:
                for k in env:
                    del os.environ[k]
            s2.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
            s2.solve_will(t(s["B"]), eps=args.eps, iters=1000)
            tunes.append(s2.tuning())
            s2.close()
            progress(f"counted forms {tunes[-1]}")
        sol.solve_will(t(s["B"]), eps=args.eps, iters=1000)
        sol.time_kernel(4, 3)
        torch.cuda.synchronize()
        os.makedirs(args.pmc_run, exist_ok=True)
        with open(os.path.join(args.pmc_run, "rows.json"), "w") as fh:
            json.dump({"rows": N, "forms": tunes}, fh)
        print(json.dumps({"rows": N, "forms": tunes}))
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
    # the dominant kernel class of a solve, from the in-solve profile (each
    # dispatch stamped by a pair of HIP events on the solver's stream)
    sol.kernel_profile(True)
    sol.solve_will(rhs[0], out, eps=args.eps, iters=1000)
    prof = sol.kernel_profile(False)["fp64"]
    t_copy = sol.time_kernel(4, 50)
    ix = "Ix%d" % sol.index_bits()
    tune = sol.tuning()
    st = tune["stream64"]
    sp, sm = ("true" if st > 0 and st & 2 else "false"), ("true" if st > 0 and st & 1 else "false")
    names = {"spmv": f"k_spmv_v2<double,4,0,{ix},{sp},false>", "pm_stage": f"k_pm_v2<double,1,0,{ix},{sm},double>",
             "pm_pair": f"k_pm2_v2<double,{ix},{sm},double>", "pm_dot": f"k_pm_v2<double,1,DOT,{ix},{sm},double>",
             "p_update": "k_pupd_vec<double>", "e_update": "k_eupd_vec<double>", "xr_update": "k_update_vec<double>"}
    cls = max((c for c in WILL_CLASS_BYTES if prof[c][0] > 0), key=lambda c: prof[c][1])
    n_l, ms_l = prof[cls]
    kname, kms, kbytes = names[cls], ms_l / n_l, WILL_CLASS_BYTES[cls] * N
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
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel": kname, "kernel_class": cls,
                     "launches": n_l, "alg_bytes_per_launch": kbytes, "avg_us": kms * 1e3,
                     "in_solve_us": {c: round(v[1] * 1e3 / v[0], 2) for c, v in prof.items() if v[0] > 0},
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


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench: {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


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


def host_info():
    """The host the CPU baseline runs on: logical CPUs, the CPUs this process
    may run on, and lscpu's model, sockets and cores."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def host_system(g):
    """The configuration's pressure system assembled on the host by the C
    oracle (the checker's restatement of set_nflag, the row numbering,
    init_matrix, cal_matrix_element and the RHS), in the reference caller's
    layout: AL(NA,NL), LL(NA,NL) column-major with NA = N + 60
    (1sub-momentum.F90:254-259), D, B length N."""
    import numpy as np
    orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    A = g.arrays()
    nc = g.ncell
    nf, nb = np.zeros(nc, np.int8), np.zeros(nc, np.int8)
    r2c, mno = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    AL = np.zeros((6, nc))
    LL = np.zeros((6, nc), np.int32)
    D, B = np.zeros(nc), np.zeros(nc)
    orc.oracle_build_system.restype = ctypes.c_int
    N = orc.oracle_build_system(g.nx, g.ny, g.nz, P(A["solid"]), P(A["F"]), P(A["fb"]), P(A["a"]), P(A["un"]),
                                P(A["dx"]), P(A["xf"]), g.stride, ctypes.c_double(g.dt), P(nf), P(nb), P(r2c),
                                P(mno), nc, P(AL), P(LL), P(D), P(B))
    NA = N + 60
    return {"N": N, "NA": NA, "AL": np.asfortranarray(AL[:, :NA].T), "LL": np.asfortranarray(LL[:, :NA].T),
            "D": D[:N].copy(), "B": B[:N].copy()}


def reference_solve(sysm, threads, eps, in_thread=True):
    """One call of the reference's own PCGPME (flang build of
    pcgpme_para.F90 in oracle/_ref; the C oracle if that was not built) on
    copies of the system, at `threads` OpenMP threads, X = 0 as the caller
    passes it (1sub-momentum.F90:559-564).  The reference's work arrays are
    automatic (stack) arrays: in_thread runs it in a thread with a stack of
    their size, else in the calling thread.  Returns (seconds, ITER, IER, X,
    kind)."""
    import numpy as np
    N, NA = sysm["N"], sysm["NA"]
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_pcgpme.so")
    kind = "reference" if os.path.exists(ref_path) else "port"
    a, l = sysm["AL"].copy(order="F"), sysm["LL"].copy(order="F")
    d, b = sysm["D"].copy(), sysm["B"].copy()
    x = np.zeros(N + 64)
    epsv = np.array([eps, 0.0])
    it, ier = ctypes.c_int(10000), ctypes.c_int(0)
    lib = None
    if kind == "reference":
        lib = ctypes.CDLL(ref_path)
        i32 = lambda v: ctypes.byref(ctypes.c_int(v))

        def call():
            # the OpenMP thread count is a per-thread setting: set it in the
            # thread that calls the reference
            lib.ref_set_threads(threads)
            lib.pcgpme_(P(a), i32(6), i32(NA), P(l), P(d), i32(N), i32(64), P(b), P(x), P(epsv),
                        ctypes.byref(ctypes.c_double(5.0)), ctypes.byref(it), ctypes.byref(ier))
    else:
        orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))

        def call():
            orc.oracle_set_threads(threads)
            orc.oracle_pcgpme(P(a), 6, NA, P(l), P(d), N, 64, P(b), P(x), P(epsv), ctypes.c_double(5.0),
                              ctypes.byref(it), ctypes.byref(ier))
    out = {}

    def body():
        t0 = time.perf_counter()
        call()
        out["t"] = time.perf_counter() - t0

    with _quiet_stdout(lib):
        if in_thread:
            old = threading.stack_size()
            threading.stack_size(max(32 << 20, 128 * (N + 64) + 32 * NA + (64 << 20)))
            th = threading.Thread(target=body)
            th.start()
            th.join()
            threading.stack_size(old)
        else:
            body()
    return out["t"], it.value, ier.value, x[:N].copy(), kind


def reference_worker(args):
    """Child process of the CPU baseline: builds the configuration's system
    on the host and times the reference's PCGPME in its main thread, so that
    the OpenMP runtime starts in this thread and binds its threads to
    consecutive cores (OMP_PROC_BIND=close).  Writes X and the timings next
    to args.reference_worker."""
    import importlib.util
    import numpy as np
    # the configuration module alone: the package would load torch, whose
    # OpenMP runtime binds this thread to one core under OMP_PROC_BIND
    spec = importlib.util.spec_from_file_location(
        "clowns_configs", os.path.join(ROOT, "2clowns-3d_amd", "clowns3d", "configs.py"))
    configs = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(configs)
    g = configs.CONFIGS[args.config]()
    # The oracle's assembly uses the GNU OpenMP runtime, which binds the
    # thread it starts in to one core under OMP_PROC_BIND; a thread of its
    # own keeps the main thread, where the reference's runtime starts, free.
    out = {}
    th = threading.Thread(target=lambda: out.setdefault("sys", host_system(g)))
    th.start()
    th.join()
    sysm = out["sys"]
    t, it, _, _, _ = reference_solve(sysm, args.cpu_threads, args.eps, in_thread=False)
    progress(f"reference warm-up at {args.cpu_threads} threads: {t:.2f} s, ITER {it}")
    times, iters, ier, x, kind = [], [], 0, None, ""
    for r in range(args.cpu_runs):
        t, it, ier, x, kind = reference_solve(sysm, args.cpu_threads, args.eps, in_thread=False)
        progress(f"reference run {r + 1}/{args.cpu_runs}: {t:.2f} s, ITER {it}, IER {ier}")
        times.append(t)
        iters.append(it)
    np.save(args.reference_worker + ".npy", x)
    with open(args.reference_worker + ".json", "w") as fh:
        json.dump({"times": times, "iters": iters, "ier": ier, "kind": kind, "N": sysm["N"]}, fh)


def cpu_baseline(config, threads, runs, eps, gpu_x):
    """The reference solve on the host cores, in a child process started
    with OMP_NUM_THREADS = threads, OMP_PROC_BIND = close, OMP_PLACES = cores
    and a stack limit the reference's automatic arrays fit in: one warm-up
    call, then the median of `runs` timed calls, each to convergence.  Also
    the reference's ITER and how far each GPU solution lies from the
    reference's X."""
    import resource
    import tempfile
    import numpy as np
    fd, base = tempfile.mkstemp(prefix="clowns_ref_")
    os.close(fd)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")

    def big_stack():
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        want = resource.RLIM_INFINITY if hard == resource.RLIM_INFINITY else hard
        resource.setrlimit(resource.RLIMIT_STACK, (want, hard))

    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--config", str(config), "--eps", repr(eps),
           "--cpu-threads", str(threads), "--cpu-runs", str(runs), "--reference-worker", base]
    r = subprocess.run(cmd, env=env, preexec_fn=big_stack, stdout=subprocess.DEVNULL)
    try:
        if r.returncode != 0:
            raise RuntimeError(f"CPU reference worker exited with status {r.returncode}")
        with open(base + ".json") as fh:
            w = json.load(fh)
        x_ref = np.load(base + ".npy")
    finally:
        for f in (base, base + ".json", base + ".npy"):
            if os.path.exists(f):
                os.remove(f)
    hi = host_info()
    rel = {k: float(np.linalg.norm(v - x_ref) / np.linalg.norm(x_ref)) for k, v in gpu_x.items()}
    kind = w["kind"]
    return {"value": float(np.median(w["times"])), "unit": "s/step", "cores": threads, "kind": kind,
            "sample": (f"the configuration's system (N={w['N']}) assembled on the host, "
                       f"{'flang build of pcgpme_para.F90' if kind == 'reference' else 'C oracle of PCGPME'} "
                       f"run to convergence (its own stopping rule, eps {eps:g}) at {threads} OpenMP threads "
                       f"(OMP_PROC_BIND=close, OMP_PLACES=cores) in a child process; median of {runs} after "
                       f"one warm-up"),
            "runs_s": w["times"], "ref_iters": w["iters"][-1], "ref_ier": w["ier"], "host": hi,
            "rel_l2_gpu_vs_ref": rel}


# Algorithmic bytes per row of each kernel class (values and vectors the
# kernel must stream once; index arrays excluded), FP64.  The FP32 classes of
# the mixed inner solve move half.  Whole iteration: 512 B/row (FP64), 256
# (FP32); per FP64 true residual of the mixed solve 216.
CLASS_BYTES = {"spmv": 72, "pm_stage": 24, "pm_pair": 32, "pm_dot": 40, "p_update": 32, "e_update": 24,
               "xr_update": 56, "residual_spmv": 80}


def kernel_names(tune, precision):
    """Names of the kernel forms the solver runs on this operator, per class."""
    if precision == "fp64":
        st = tune["stream64"]
        pm, sp = ("true" if st > 0 and st & 1 else "false"), ("true" if st > 0 and st & 2 else "false")
        ix = "Ix%d" % tune["index64"]
        pat = "true" if tune.get("spmv32", -1) >= 0 and tune["spmv32"] & 4 else "false"
        names = {"spmv": f"k_spmv_v2<double,6,0,{ix},{sp},{pat}>", "pm_stage": f"k_pm_v2<double,1,0,{ix},{pm},double>",
                 "pm_pair": f"k_pm2_v2<double,{ix},{pm},double>", "pm_dot": f"k_pm_v2<double,2,DOT,{ix},{pm},double>",
                 "p_update": "k_pupd_vec<double>", "e_update": "k_eupd_vec<double>",
                 "xr_update": "k_update_vec<double>", "residual_spmv": f"k_spmv_v2<double,6,3,{ix},{sp},{pat}>"}
        if tune.get("pmpat64") == 1:
            names["pm_stage"] = f"k_pm_p2<1,0,{ix},{pm}>"
            names["pm_dot"] = f"k_pm_p2<2,DOT,{ix},{pm}>"
        return names
    st = tune["stream32"]
    pm, sp = ("true" if st > 0 and st & 1 else "false"), ("true" if st > 0 and st & 2 else "false")
    ix = "Ix%d" % tune["index32"]
    codes = "unsigned char" if ix == "Ix8" else "short"
    c = "_Float16" if os.environ.get("CLOWNS_PM_HALF", "1") != "0" else "float"
    form = tune.get("spmv32", 1)
    pat = "true" if form >= 0 and form & 4 else "false"
    if form >= 0 and (form & 3) == 2:
        spmv = f"k_spmv_rows32_p2<{ix},{codes},{sp},{pat}>"
    elif form >= 0 and (form & 3) == 0:
        spmv = f"k_spmv_v2<float,6,0,{ix},{sp},{pat}>"
    else:
        spmv = f"k_spmv_rows32<{ix},{codes},{sp},{pat}>"
    return {"spmv": spmv, "pm_stage": f"k_pm_v4<1,0,{ix},{pm},{c}>",
            "pm_pair": f"k_pm2_v2<float,{ix},{pm},{c}>", "pm_dot": f"k_pm_v4<2,DOT,{ix},{pm},{c}>",
            "p_update": "k_pupd_vec<float>", "e_update": "k_eupd_vec<float>", "xr_update": "k_update_vec<float>"}


def dominant(prof, precision, N, tune, traffic):
    """The kernel class with the most in-solve time: its name, launches,
    average duration and algorithmic rate."""
    scale = 1.0 if precision == "fp64" else 0.5
    classes = {c: v for c, v in prof[precision].items() if v[0] > 0 and c in CLASS_BYTES and c != "residual_spmv"}
    cls = max(classes, key=lambda c: classes[c][1])
    n, ms = classes[cls]
    avg_us = ms * 1e3 / n
    alg = CLASS_BYTES[cls] * scale * N
    name = kernel_names(tune, precision)[cls]
    achieved = alg / (avg_us * 1e-6) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": achieved / PEAK_HBM_GBS, "traffic": (traffic or {}).get(name), "kernel": name,
           "kernel_class": cls, "launches": n, "avg_us": avg_us, "alg_bytes_per_launch": alg,
           "alg_bytes_per_row": CLASS_BYTES[cls] * scale,
           "in_solve_us": {c: round(v[1] * 1e3 / v[0], 2) for c, v in prof[precision].items() if v[0] > 0},
           "launches_by_class": {c: v[0] for c, v in prof[precision].items() if v[0] > 0}}
    return out


def traffic_from_pmc(fetch_dir, write_dir, rows, out_path):
    """HBM traffic per launch from two rocprofv3 --pmc runs of
    `bench.py --pmc-run` (FETCH_SIZE in one, WRITE_SIZE in the other; KiB per
    dispatch), each scaled by what the counter reads for the plain copy
    kernel of 8 N bytes read and 8 N written that the run launches.  A
    kernel's figure is the median over its dispatches that moved data (those
    queued past convergence read only the run's scalars)."""
    import csv
    import glob
    import statistics

    def read(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = _short_name(row["Kernel_Name"])
                    vals.setdefault(name, []).append(float(row["Counter_Value"]))
        return vals

    fv, wv = read(fetch_dir, "FETCH_SIZE"), read(write_dir, "WRITE_SIZE")
    copy = "k_copy<double>"
    ff = statistics.median(fv[copy]) * 1024 / (8 * rows)
    wf = statistics.median(wv[copy]) * 1024 / (8 * rows)
    out = {"rows": rows, "fetch_factor": ff, "write_factor": wf}
    for name in sorted(set(fv) & set(wv)):
        if not name.startswith("k_"):
            continue
        f = [v for v in fv[name] if v > 0.01 * max(fv[name])]
        w = [v for v in wv[name] if v > 0.01 * max(wv[name])] or [0.0]
        out[name] = (statistics.median(f) * 1024 / ff + statistics.median(w) * 1024 / wf)
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    return out


def _demangle_args(m):
    """Template arguments of a mangled clowns:: kernel name that the
    profiler left mangled (its demangler stops at _Float16, 'DF16_'):
    '_ZN6clowns7k_pm_v4ILi1ELi0ENS_3Ix8ELb0EDF16_EE...' -> 'k_pm_v4<1,0,Ix8,false,_Float16>'."""
    import re
    head = re.match(r"_ZN6clowns(\d+)", m)
    if not head:
        return m
    n = int(head.group(1))
    pos = head.end()
    name = m[pos:pos + n]
    pos += n
    if pos >= len(m) or m[pos] != "I":
        return name
    pos += 1
    args = []
    while pos < len(m) and m[pos] != "E":
        t = m[pos:]
        if t.startswith("Li"):
            e = t.index("E")
            args.append(t[2:e])
            pos += e + 1
        elif t.startswith("Lb"):
            args.append("true" if t[2] == "1" else "false")
            pos += 4
        elif t.startswith("NS_"):
            mm = re.match(r"NS_(\d+)", t)
            k = int(mm.group(1))
            args.append(t[mm.end():mm.end() + k])
            pos += mm.end() + k + 1
        elif t.startswith("DF16_"):
            args.append("_Float16")
            pos += 5
        elif t[0] in "fdish":
            args.append({"f": "float", "d": "double", "i": "int", "s": "short", "h": "unsigned char"}[t[0]])
            pos += 1
        else:
            return m
    return name + "<" + ",".join(args) + ">"


def _short_name(k):
    """rocprof's demangled kernel name in the form kernel_names uses:
    'void clowns::k_pm_v2<double, 1, 0, clowns::Ix8, true, double>(...)' ->
    'k_pm_v2<double,1,0,Ix8,true,double>' (the dot-carrying stages' DOT
    argument becomes the word DOT)."""
    if k.startswith("_ZN6clowns"):
        k = _demangle_args(k)
    k = k.split("(")[0].replace("void ", "").replace("clowns::", "").replace(" ", "").strip()
    for pre in ("k_pm_v2<double,2,", "k_pm_v4<2,", "k_pm_v4<1,", "k_pm_p2<2,", "k_pm_p2<1,"):
        if k.startswith(pre):
            rest = k[len(pre):]
            dot, _, tail = rest.partition(",")
            if pre.endswith("<1,") and dot == "0":
                return k
            return pre + "DOT," + tail
    if k.startswith("k_pm_v2<double,1,") and not k.startswith("k_pm_v2<double,1,0,"):
        rest = k[len("k_pm_v2<double,1,"):]
        return "k_pm_v2<double,1,DOT," + rest.partition(",")[2]
    return k.replace("unsignedchar", "unsigned char")


def trace_summary(csv_path, top=24):
    """Markdown table of a rocprofv3 --kernel-trace CSV: per kernel, all
    dispatches and the working ones -- those longer than a quarter of the
    kernel's 90th-percentile duration (launches queued past a solve's
    convergence return at once) -- with the working launches' average."""
    import csv
    durs = {}
    for row in csv.DictReader(open(csv_path)):
        name = _short_name(row["Kernel_Name"])
        durs.setdefault(name, []).append((int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3)
    out = []
    for name, d in durs.items():
        d_sorted = sorted(d)
        p90 = d_sorted[int(0.9 * (len(d_sorted) - 1))]
        work = [v for v in d if v > 0.25 * p90]
        out.append((sum(work), name, len(d), len(work), sum(work) / max(1, len(work))))
    out.sort(reverse=True)
    total = sum(o[0] for o in out)
    lines = ["| kernel | launches | working launches | avg us (working) | total ms (working) | % |",
             "|---|---|---|---|---|---|"]
    for tot, name, n, nw, avg in out[:top]:
        lines.append(f"| `{name}` | {n} | {nw} | {avg:.2f} | {tot / 1e3:.2f} | {100 * tot / total:.1f} |")
    return "\n".join(lines)


def boundary_mode(sysm, eps, reps=2):
    """The drop-in clowns_pcgpme with the caller's host arrays (pageable
    numpy memory, as a Fortran caller holds them): matrix, right-hand side
    and solution cross PCIe on every call.  The last of `reps` calls is
    reported (the first carries the library solver's one-time form tuning)."""
    import numpy as np
    import clowns3d
    N, NA = sysm["N"], sysm["NA"]
    t = None
    for _ in range(reps):
        a, l = sysm["AL"].copy(order="F"), sysm["LL"].copy(order="F")
        d, b = sysm["D"].copy(), sysm["B"].copy()
        x = np.zeros(N + 64)
        epsv = np.array([eps, 0.0])
        t0 = time.perf_counter()
        it, ier = clowns3d.pcgpme(a, 6, NA, l, d, N, 64, b, x, epsv, 5.0, 10000)
        t = time.perf_counter() - t0
    mb = (a.nbytes + l.nbytes + d.nbytes + b.nbytes + 8 * (N + 64)) / 1e6
    return {"s_per_solve": t, "iters": it, "ier": ier, "host_MB_each_way": mb}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the rows into z-slabs, one per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.reference_worker:
        reference_worker(args)
        return
    if args.trace_summary:
        print(trace_summary(args.trace_summary))
        return
    if args.traffic:
        fetch_dir, write_dir = args.traffic
        with open(os.path.join(fetch_dir, "rows.json")) as fh:
            rows = json.load(fh)["rows"]
        name = f"traffic_will{args.will.lower()}.json" if args.will else f"traffic_config{args.config}.json"
        out = traffic_from_pmc(fetch_dir, write_dir, rows, os.path.join(ROOT, "profiles", name))
        print(json.dumps(out))
        return
    if args.will:
        bench_will(args)
        return
    if args.vof:
        bench_vof(args)
        return
    import numpy as np
    import torch
    from clowns3d import DeviceGrid, Solver, configs

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    g = configs.CONFIGS[args.config]()
    dg = DeviceGrid(g)
    N = dg.N
    progress(f"configuration {args.config}: {N} rows classified and resident")
    sol = Solver(0)
    b = torch.empty(N, dtype=torch.float64, device=dev)
    x = torch.zeros(N, dtype=torch.float64, device=dev)

    def step(mixed):
        dg.assemble_into(sol, b=b)
        x.zero_()
        _, info = sol.solve(b, x, eps=args.eps, maxiter=10000, mixed=mixed)
        return info

    if args.pmc_run:
        # an assembled FP64 solve and a mixed solve in each of the kernel
        # forms the solver may pick here (streamed or cached PM stages and
        # SpMV, 8- or 16-bit index codes, single or fused PM stages, FP64 PM
        # stages with or without the pattern table), so the traffic of
        # whichever form a timed run picks was counted inside a solve; then
        # the copy kernel the counters are scaled by
        tunes = []
        for st, ll8, fuse, pat in itertools.product(("0", "1", "2", "3"), ("0", "1"), ("0", "1"), ("0", "1")):
            env = {"CLOWNS_STREAM": st, "CLOWNS_LL8": ll8, "CLOWNS_FUSE_PM": fuse, "CLOWNS_PM_PAT": pat}
            os.environ.update(env)
            try:
                s2 = Solver(0)
            finally:
                for k in env:
                    del os.environ[k]
            for mixed in (False, True):
                dg.assemble_into(s2, b=b)
                x.zero_()
                s2.solve(b, x, eps=args.eps, maxiter=10000, mixed=mixed)
            tunes.append(s2.tuning())
            s2.close()
            progress(f"counted forms {tunes[-1]}")
        dg.assemble_into(sol, b=b)
        sol.time_kernel(4, 3)
        torch.cuda.synchronize()
        os.makedirs(args.pmc_run, exist_ok=True)
        with open(os.path.join(args.pmc_run, "rows.json"), "w") as fh:
            json.dump({"rows": N, "forms": tunes}, fh)
        print(json.dumps({"rows": N, "forms": tunes}))
        return

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
    progress(f"FP64: {t64 * 1e3:.1f} ms/step, {infos64[-1]['iters']} iterations")
    x64 = x.clone()
    it64 = infos64[-1]["iters"]
    assert all(i["ier"] == 0 for i in infos64), infos64

    # accuracy reference for both forms: a tight FP64 solve (stagnation exit)
    dg.assemble_into(sol, b=b)
    xt = torch.zeros_like(x)
    sol.solve(b, xt, eps=2e-15, maxiter=20000)
    err64 = ((x64 - xt).norm() / xt.norm()).item()
    mixed = None
    x32 = None
    if not args.no_mixed:
        t32, infos32 = run(True)
        progress(f"mixed: {t32 * 1e3:.1f} ms/step, {infos32[-1]['iters']} inner iterations")
        x32 = x.clone()
        i32 = infos32[-1]
        mixed_bytes = N * (ALG_BYTES["setup_cold"] + i32["iters"] * ALG_BYTES["iteration32"] +
                           i32["refinements"] * ALG_BYTES["refine64"])
        mixed_gbs = mixed_bytes / (i32["ms"] * 1e-3) / 1e9
        mixed = {"value": t32, "ms_per_step": t32 * 1e3, "inner_iters": i32["iters"],
                 "refinements": i32["refinements"], "err": i32["err"], "ier": i32["ier"],
                 "solve_ms_hip_events": i32["ms"], "solve_alg_GBps": mixed_gbs,
                 "solve_alg_frac": mixed_gbs / PEAK_HBM_GBS,
                 "rel_l2_vs_tight_fp64": ((x32 - xt).norm() / xt.norm()).item(),
                 "rel_l2_vs_fp64": ((x32 - x64).norm() / x64.norm()).item(), "speedup_vs_fp64": t64 / t32}

    # In-solve kernel durations: HIP events around every launch of one more
    # step of each form (identical to the timed ones), on the solver's stream.
    sol.kernel_profile(True)
    step(False)
    if not args.no_mixed:
        step(True)
    prof = sol.kernel_profile(False)
    progress("in-solve kernel profile taken")
    tune = sol.tuning()
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"traffic_config{args.config}.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            traffic = json.load(fh)
    roofline = dominant(prof, "fp64", N, tune, traffic)
    roofline["copy_GBps"] = 16 * N / (sol.time_kernel(4, 20) * 1e-3) / 1e9
    if mixed:
        mixed["roofline"] = dominant(prof, "fp32", N, tune, traffic)
    solve_ms = infos64[-1]["ms"]
    solve_gbs = N * (ALG_BYTES["setup_cold"] + it64 * ALG_BYTES["iteration"]) / (solve_ms * 1e-3) / 1e9

    cpu, boundary = None, None
    if not (args.no_cpu and args.no_boundary):
        sysm = host_system(g)
        progress(f"host system assembled ({sysm['N']} rows)")
        assert sysm["N"] == N, (sysm["N"], N)
        # the host system is the device's, bit for bit
        bh = b.cpu().numpy()
        system_same = bool(np.array_equal(bh, sysm["B"]))
        if not args.no_boundary:
            boundary = boundary_mode(sysm, args.eps)
            progress(f"drop-in with host arrays: {boundary['s_per_solve'] * 1e3:.1f} ms")
            boundary["resident_s_per_solve"] = t64
            boundary["ratio_to_resident"] = boundary["s_per_solve"] / t64
        del sysm
        if not args.no_cpu:
            # the host's cores, up to the 16 a one-GPU host grants a process
            threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
            gx = {"fp64": x64.cpu().numpy()}
            if x32 is not None:
                gx["mixed"] = x32.cpu().numpy()
            gx["tight_fp64"] = xt.cpu().numpy()
            cpu = cpu_baseline(args.config, threads, args.cpu_runs, args.eps, gx)
            cpu["gpu_iters"] = it64
            cpu["host_rhs_bitexact"] = system_same

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
        "boundary": boundary,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = cpu["value"] / t64
    print(json.dumps(out))


if __name__ == "__main__":
    main()
