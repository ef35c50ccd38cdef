"""Host-side mirror of the reference's PCGPME interface over the C-ABI.

``pcgpme`` keeps the argument list and in/out semantics of
``subroutine PCGPME(AL,NL,NA,LL,D,N,N2,B,X,EPS,PARM,ITER,IER)``
(src2LOWNS/pcgpme_para.F90:1-58, 250-257) on numpy arrays laid out as the
Fortran caller holds them: AL/LL column-major (NA, NL), LL 1-based.  Like the
reference it mutates AL, LL, D, B, X and EPS in place and returns
(ITER, IER).  ``pcgpme_will`` does the same for the 2DH dispersion solve
``PCGPME_WILL(AL,NL,NA,LL,D,N,N2,B,EPS,PARM,ITER)`` (src2LOWNS/pcgpme_2DH.f90:1-41):
only B changes (right-hand side in, filtered solution out).  ``Solver`` is
the device-resident form on torch tensors.
"""
import ctypes

import numpy as np

from . import _lib as L


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check_host(name, a, dtype, shape=None, fortran=False):
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise TypeError(f"{name} must be a numpy {np.dtype(dtype).name} array")
    if shape is not None and a.shape != shape:
        raise ValueError(f"{name} has shape {a.shape}, expected {shape}")
    if fortran and not a.flags.f_contiguous:
        raise ValueError(f"{name} must be Fortran-ordered (column-major)")
    if not fortran and not a.flags.c_contiguous:
        raise ValueError(f"{name} must be contiguous")


def pcgpme(AL, NL, NA, LL, D, N, N2, B, X, EPS, PARM, ITER):
    """Drop-in of PCGPME through ``clowns_pcgpme``.  Returns (ITER, IER)."""
    _check_host("AL", AL, np.float64, (NA, NL), fortran=True)
    _check_host("LL", LL, np.int32, (NA, NL), fortran=True)
    _check_host("D", D, np.float64, (N,))
    _check_host("B", B, np.float64, (N,))
    _check_host("X", X, np.float64, (N + N2,))
    _check_host("EPS", EPS, np.float64, (2,))
    it = ctypes.c_int32(int(ITER))
    ier = ctypes.c_int32(0)
    i32 = lambda v: ctypes.byref(ctypes.c_int32(int(v)))
    L.lib().clowns_pcgpme(_ptr(AL), i32(NL), i32(NA), _ptr(LL), _ptr(D), i32(N), i32(N2),
                          _ptr(B), _ptr(X), _ptr(EPS), ctypes.byref(ctypes.c_double(PARM)),
                          ctypes.byref(it), ctypes.byref(ier))
    return it.value, ier.value


def pcgpme_will(AL, NL, NA, LL, D, N, N2, B, EPS, PARM, ITER):
    """Drop-in of PCGPME_WILL through ``clowns_pcgpme_will``.  B is
    overwritten with the reference's output; returns IER (0, 3000 or a
    CLOWNS_E* code -- the reference itself has no status argument)."""
    _check_host("AL", AL, np.float64, (NA, NL), fortran=True)
    _check_host("LL", LL, np.int32, (NA, NL), fortran=True)
    _check_host("D", D, np.float64, (N,))
    _check_host("B", B, np.float64, (N,))
    ier = ctypes.c_int32(0)
    i32 = lambda v: ctypes.byref(ctypes.c_int32(int(v)))
    f64 = lambda v: ctypes.byref(ctypes.c_double(float(v)))
    L.lib().clowns_pcgpme_will(_ptr(AL), i32(NL), i32(NA), _ptr(LL), _ptr(D), i32(N), i32(N2),
                               _ptr(B), f64(EPS), f64(PARM), i32(ITER), ctypes.byref(ier))
    return ier.value


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Solver:
    """Device-resident PCGPME: the scaled/sorted operator stays in HBM.

    Tensors are torch CUDA tensors on the solver's device.  Work runs on the
    current torch stream at construction (``stream=None``) so torch events
    and later torch work order correctly around it.
    """

    def __init__(self, device=0, stream=None):
        import torch
        self._torch = torch
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        h = ctypes.c_void_p()
        L.check(L.lib().clowns_solver_create(ctypes.byref(h), int(device), ctypes.c_void_p(stream)),
                "clowns_solver_create")
        self._h = h
        self.device = device
        self.N = 0
        self.lda = 0
        self.NL = 0
        self.version = 0  # bumped whenever the resident operator changes

    def close(self):
        if self._h:
            L.lib().clowns_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_matrix(self, AL, LL, D, parm=5.0):
        """AL (NL, lda) float64, LL (NL, lda) int32 1-based, D (N,) -- i.e.
        the reference's column-major AL(NA,NL) viewed row-major as (NL, NA)."""
        torch = self._torch
        NL, lda = AL.shape
        N = D.shape[0]
        for name, t, dt in (("AL", AL, torch.float64), ("LL", LL, torch.int32), ("D", D, torch.float64)):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous CUDA {dt} tensor")
        if LL.shape != AL.shape:
            raise ValueError("LL and AL shapes differ")
        L.check(L.lib().clowns_solver_set_matrix(self._h, _dptr(AL), NL, lda, _dptr(LL), _dptr(D), N,
                                                 float(parm)), "clowns_solver_set_matrix")
        self.N, self.lda, self.NL = N, lda, NL
        self.version += 1

    def solve(self, b, x=None, eps=1e-10, maxiter=10000, mixed=False, warm=False):
        torch = self._torch
        if x is None:
            x = torch.zeros_like(b)
        flags = (L.CLOWNS_MIXED if mixed else 0) | (L.CLOWNS_WARM_START if warm else 0)
        info = L.Info()
        L.check(L.lib().clowns_solver_solve(self._h, _dptr(b), _dptr(x), float(eps), int(maxiter), flags,
                                            ctypes.byref(info)), "clowns_solver_solve")
        return x, info.as_dict()

    def solve_noalloc(self, b, x, eps=1e-10, maxiter=10000, mixed=False, warm=False):
        """solve() through clowns_solver_solve_noalloc: no allocation in the
        call (work memory in the library's fixed array; at most 524288 rows;
        mixed needs one solve(mixed=True) first).  Returns (x, info);
        raises ClownsError on a rejected call."""
        flags = (L.CLOWNS_MIXED if mixed else 0) | (L.CLOWNS_WARM_START if warm else 0)
        info = L.Info()
        L.check(L.lib().clowns_solver_solve_noalloc(self._h, _dptr(b), _dptr(x), float(eps), int(maxiter), flags,
                                                    ctypes.byref(info)), "clowns_solver_solve_noalloc")
        return x, info.as_dict()

    def index_bits(self):
        """8 or 16 if the resident operator's kernels read 8- or 16-bit coded
        neighbour indices, else 32."""
        return L.lib().clowns_solver_index_bits(self._h)

    def set_block_order(self, order=None):
        """Order of the 512-row blocks in the resident operator's SpMV and PM
        kernels (clowns_solver_set_block_order): an int32 CUDA tensor holding
        a permutation of 0..G-1, G = ceil(N/512) rounded up to 8; None for
        the natural order.  Results are bit-identical either way."""
        if order is None:
            L.check(L.lib().clowns_solver_set_block_order(self._h, None, 0), "clowns_solver_set_block_order")
            return
        order = order.contiguous()
        L.check(L.lib().clowns_solver_set_block_order(self._h, _dptr(order), int(order.numel())),
                "clowns_solver_set_block_order")

    def tuning(self):
        """Kernel forms chosen for the resident operator (clowns_solver_tuning):
        dict fuse64, fuse32 (1 fused PM pairs), stream64, stream32 (bit 0
        streamed PM stages, bit 1 streamed SpMV), index64, index32 (bits of
        the neighbour-index codes: 8, 16 or 32), pmpat64 (1: FP64 PM stages
        with the row-pattern table), spmv32 (FP32 SpMV: 0 column-major, 1
        row-major copies, 2 row-major in two chunks a wavefront, plus 4 with
        the row-pattern table); -1 where undecided."""
        out = np.zeros(8, np.int32)
        L.check(L.lib().clowns_solver_tuning(self._h, ctypes.c_void_p(out.ctypes.data)), "clowns_solver_tuning")
        return dict(zip(("fuse64", "fuse32", "stream64", "stream32", "index64", "index32", "pmpat64", "spmv32"),
                        (int(v) for v in out)))

    def solve_will(self, b, out=None, eps=1e-10, iters=1000):
        """PCGPME_WILL on the resident operator (set_matrix): zero start, EPS
        and ITER as given, then the reference's output pass.  Returns
        (out, info) with info["filter_rounds"] the device rounds of that pass
        (-1: it ran on the host)."""
        torch = self._torch
        if out is None:
            out = torch.empty_like(b)
        info = L.Info()
        rounds = ctypes.c_int32(0)
        L.check(L.lib().clowns_solver_solve_will(self._h, _dptr(b), _dptr(out), float(eps), int(iters),
                                                 ctypes.byref(info), ctypes.byref(rounds)),
                "clowns_solver_solve_will")
        d = info.as_dict()
        d["filter_rounds"] = rounds.value
        return out, d

    def will_filter(self, LL, X, R):
        """PCGPME_WILL's output pass alone: LL (NL, lda) int32 caller order,
        X and R (N,) float64.  Returns (X_filtered, B, rounds)."""
        torch = self._torch
        NL, lda = LL.shape
        N = X.shape[0]
        for name, t, dt in (("LL", LL, torch.int32), ("X", X, torch.float64), ("R", R, torch.float64)):
            if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous CUDA {dt} tensor")
        xo = torch.empty_like(X)
        bo = torch.empty_like(R)
        rounds = ctypes.c_int32(0)
        L.check(L.lib().clowns_will_filter(self._h, _dptr(LL), NL, lda, N, _dptr(X), _dptr(R), _dptr(xo),
                                           _dptr(bo), ctypes.byref(rounds)), "clowns_will_filter")
        return xo, bo, rounds.value

    def solve_transpose(self, g, lam=None, eps=1e-10, maxiter=10000, warm=False):
        torch = self._torch
        if lam is None:
            lam = torch.zeros_like(g)
        info = L.Info()
        L.check(L.lib().clowns_solver_solve_transpose(self._h, _dptr(g), _dptr(lam), float(eps), int(maxiter),
                                                      L.CLOWNS_WARM_START if warm else 0, ctypes.byref(info)),
                "clowns_solver_solve_transpose")
        return lam, info.as_dict()

    def grad_coefficients(self, lam, x):
        torch = self._torch
        gAL = torch.zeros((self.NL, self.lda), dtype=torch.float64, device=lam.device)
        gD = torch.zeros(self.N, dtype=torch.float64, device=lam.device)
        L.check(L.lib().clowns_grad_coefficients(self._h, _dptr(lam), _dptr(x), _dptr(gAL), _dptr(gD)),
                "clowns_grad_coefficients")
        return gAL, gD

    def get_scaled(self, residual=False):
        """The resident scaled/sorted operator (AL, LL, D), the row scales W
        and, with residual=True (after a solve), the final scaled residual."""
        torch = self._torch
        dev = f"cuda:{self.device}"
        AL = torch.empty((self.NL, self.lda), dtype=torch.float64, device=dev)
        LL = torch.empty((self.NL, self.lda), dtype=torch.int32, device=dev)
        D = torch.empty(self.N, dtype=torch.float64, device=dev)
        W = torch.empty(self.N, dtype=torch.float64, device=dev)
        B = torch.empty(self.N, dtype=torch.float64, device=dev) if residual else None
        L.check(L.lib().clowns_solver_get_scaled(self._h, _dptr(AL), _dptr(LL), _dptr(D), _dptr(W),
                                                 _dptr(B) if residual else None),
                "clowns_solver_get_scaled")
        return AL, LL, D, W, B

    KERNEL_CLASSES = ("spmv", "pm_stage", "pm_pair", "pm_dot", "p_update", "e_update", "xr_update", "finalize",
                      "residual_spmv")

    def kernel_profile(self, enable):
        """In-solve kernel durations (clowns_solver_kernel_profile).
        kernel_profile(True) arms the solver; kernel_profile(False) disarms
        it and returns {"fp64": {class: (launches, total_ms)}, "fp32": ...}
        over the solves in between (launches that did work only)."""
        if enable:
            L.check(L.lib().clowns_solver_kernel_profile(self._h, 1, None, None), "clowns_solver_kernel_profile")
            return None
        n = len(self.KERNEL_CLASSES)
        ms = np.zeros(2 * n)
        cnt = np.zeros(2 * n, np.int32)
        L.check(L.lib().clowns_solver_kernel_profile(self._h, 0, ctypes.c_void_p(ms.ctypes.data),
                                                     ctypes.c_void_p(cnt.ctypes.data)), "clowns_solver_kernel_profile")
        out = {}
        for p, off in (("fp64", 0), ("fp32", n)):
            out[p] = {c: (int(cnt[off + i]), float(ms[off + i])) for i, c in enumerate(self.KERNEL_CLASSES)}
        return out

    def patterns(self):
        """(table rows, matrix rows on them) of the resident operator's row
        patterns (clowns_solver_patterns); (0, 0) when not used."""
        n, r = ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().clowns_solver_patterns(self._h, ctypes.byref(n), ctypes.byref(r)), "clowns_solver_patterns")
        return n.value, r.value

    def time_kernel(self, kernel, reps=50):
        ms = ctypes.c_double()
        L.check(L.lib().clowns_solver_time_kernel(self._h, int(kernel), int(reps), ctypes.byref(ms)),
                "clowns_solver_time_kernel")
        return ms.value
