"""Shared test plumbing: the `gpu` marker, the oracle and reference loaders.

The oracle (oracle/liboracle.so, the plain-C restatement) and the flang build
of the reference (oracle/_ref/*.so) are checkers only.
"""
import ctypes
import os
import sys
import threading

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "2clowns-3d_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PCG_SO = os.path.join(ROOT, "oracle", "_ref", "libref_pcgpme.so")
REF_ASM_SO = os.path.join(ROOT, "oracle", "_ref", "libref_assembly.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


P = ctypes.c_void_p


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            pytest.fail(f"{ORACLE_SO} missing: run make -C oracle (or __graft_entry__.build())")
        _oracle = ctypes.CDLL(ORACLE_SO)
    return _oracle


_ref = None


def ref_pcgpme_lib():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PCG_SO):
            pytest.skip("flang build of the reference (oracle/_ref) not present")
        _ref = ctypes.CDLL(REF_PCG_SO)
    return _ref


def run_big_stack(fn, stack_bytes):
    """Run fn() in a thread with a large stack: the reference's automatic
    arrays (pcgpme_para.F90:41-49) live on the stack under flang."""
    out = {}

    def body():
        try:
            out["v"] = fn()
        except BaseException as e:  # pragma: no cover
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(max(int(stack_bytes), 32 << 20))
    try:
        th = threading.Thread(target=body)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("v")


def copy_system(s):
    return dict(AL=np.asfortranarray(s["AL"].copy()), LL=np.asfortranarray(s["LL"].copy()),
                D=s["D"].copy(), B=s["B"].copy(), N=s["N"], NA=s["NA"], NL=s["NL"])


def call_oracle(s, iters=10000, eps=1e-10, nthr=1, parm=5.0, x0=None):
    c = copy_system(s)
    N = c["N"]
    X = np.zeros(N + 64) if x0 is None else np.concatenate([x0, np.zeros(64)])
    EPS = np.array([eps, 0.0])
    it = ctypes.c_int(iters)
    ier = ctypes.c_int(-1)
    o = oracle()
    o.oracle_set_threads(nthr)
    rc = o.oracle_pcgpme(ptr(c["AL"]), c["NL"], c["NA"], ptr(c["LL"]), ptr(c["D"]), N, 64, ptr(c["B"]),
                         ptr(X), ptr(EPS), ctypes.c_double(parm), ctypes.byref(it), ctypes.byref(ier))
    c.update(X=X, EPS=EPS, ITER=it.value, IER=ier.value, rc=rc)
    return c


def call_ref(s, iters=10000, eps=1e-10, nthr=1, parm=5.0, x0=None):
    lib = ref_pcgpme_lib()
    c = copy_system(s)
    N = c["N"]
    X = np.zeros(N + 64) if x0 is None else np.concatenate([x0, np.zeros(64)])
    EPS = np.array([eps, 0.0])
    it = ctypes.c_int(iters)
    ier = ctypes.c_int(-1)
    i32 = lambda v: ctypes.byref(ctypes.c_int(v))
    # the OpenMP thread count is a per-thread setting: set it in the thread
    # that calls the reference
    run_big_stack(lambda: (lib.ref_set_threads(nthr), lib.pcgpme_(ptr(c["AL"]), i32(c["NL"]), i32(c["NA"]), ptr(c["LL"]), ptr(c["D"]),
                                      i32(N), i32(64), ptr(c["B"]), ptr(X), ptr(EPS),
                                      ctypes.byref(ctypes.c_double(parm)), ctypes.byref(it), ctypes.byref(ier))),
                  120 * (N + 64) + 64 * c["NA"] + (64 << 20))
    c.update(X=X, EPS=EPS, ITER=it.value, IER=ier.value)
    return c


def have_ref():
    return os.path.exists(REF_PCG_SO)


def host_config_system(g):
    """A configs.Grid's pressure system assembled on the host by the C oracle
    (set_nflag, row numbering, init_matrix, cal_matrix_element, RHS), in the
    reference caller's layout: AL(NA,NL), LL(NA,NL) column-major, NA = N + 60
    (src2LOWNS/1sub-momentum.F90:254-259)."""
    A = g.arrays()
    nc = g.ncell
    nf, nb = np.zeros(nc, np.int8), np.zeros(nc, np.int8)
    r2c, mno = np.zeros(nc, np.int32), np.zeros(nc, np.int32)
    AL = np.zeros((6, nc))
    LL = np.zeros((6, nc), np.int32)
    D, B = np.zeros(nc), np.zeros(nc)
    o = oracle()
    o.oracle_build_system.restype = ctypes.c_int
    N = o.oracle_build_system(g.nx, g.ny, g.nz, ptr(A["solid"]), ptr(A["F"]), ptr(A["fb"]), ptr(A["a"]),
                              ptr(A["un"]), ptr(A["dx"]), ptr(A["xf"]), g.stride, ctypes.c_double(g.dt), ptr(nf),
                              ptr(nb), ptr(r2c), ptr(mno), nc, ptr(AL), ptr(LL), ptr(D), ptr(B))
    NA = N + 60
    return dict(AL=np.asfortranarray(AL[:, :NA].T), LL=np.asfortranarray(LL[:, :NA].T), D=D[:N].copy(),
                B=B[:N].copy(), N=N, NA=NA, NL=6)


def dense_matrix(s):
    """Dense A (unscaled) from the reference-layout arrays."""
    N, NL = s["N"], s["NL"]
    A = np.zeros((N, N))
    A[np.arange(N), np.arange(N)] = s["D"][:N]
    for k in range(NL):
        col = s["LL"][:N, k]
        m = col > 0
        np.add.at(A, (np.arange(N)[m], col[m] - 1), s["AL"][:N, k][m])
    return A


def apply_A(s, x):
    """y = A x (unscaled), sparse, from the reference-layout arrays."""
    N, NL = s["N"], s["NL"]
    y = s["D"][:N] * x[:N]
    for k in range(NL):
        col = s["LL"][:N, k]
        m = col > 0
        y[m] += s["AL"][:N, k][m] * x[col[m] - 1]
    return y


def call_oracle_will(s, iters=1000, eps=1e-10, nthr=1, parm=5.0):
    """oracle_pcgpme_will on copies of s; returns dict(B, X_raw, X, IER, K,
    AL/LL/D as left after the call)."""
    c = copy_system(s)
    N = c["N"]
    xr, xo = np.zeros(N), np.zeros(N)
    ier, k = ctypes.c_int(-1), ctypes.c_int(-1)
    o = oracle()
    o.oracle_set_threads(nthr)
    rc = o.oracle_pcgpme_will(ptr(c["AL"]), c["NL"], c["NA"], ptr(c["LL"]), ptr(c["D"]), N, 64, ptr(c["B"]),
                              ctypes.c_double(eps), ctypes.c_double(parm), int(iters), ctypes.byref(ier),
                              ctypes.byref(k), ptr(xr), ptr(xo))
    c.update(X_raw=xr, X=xo, IER=ier.value, K=k.value, rc=rc)
    return c


def call_ref_will(s, iters=1000, eps=1e-10, nthr=1, parm=5.0):
    """The flang build of PCGPME_WILL (src2LOWNS/pcgpme_2DH.f90) on copies
    of s; returns dict(B, AL/LL/D as left after the call)."""
    lib = ref_pcgpme_lib()
    c = copy_system(s)
    N = c["N"]
    i32 = lambda v: ctypes.byref(ctypes.c_int(v))
    f64 = lambda v: ctypes.byref(ctypes.c_double(v))
    # PCGPME_WILL takes its chunking from omp_get_num_threads()
    # (pcgpme_2DH.f90:55-57): set the count in the thread that calls it
    run_big_stack(lambda: (lib.ref_set_threads(nthr), lib.pcgpme_will_(ptr(c["AL"]), i32(c["NL"]), i32(c["NA"]), ptr(c["LL"]), ptr(c["D"]),
                                           i32(N), i32(64), ptr(c["B"]), f64(eps), f64(parm), i32(iters))),
                  200 * (N + 64) + 64 * c["NA"] + (64 << 20))
    return c
