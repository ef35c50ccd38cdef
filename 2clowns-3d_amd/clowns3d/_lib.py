"""Loader of the in-tree gfx950 solver library (``lib/libclowns_pcg.so``).

The library is the product: there is no CPU fallback.  If it has not been
built, importing the bindings raises at once.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
LIB_PATH = os.environ.get("CLOWNS_LIB") or os.path.join(PKG_ROOT, "lib", "libclowns_pcg.so")

CLOWNS_OK = 0
CLOWNS_NOCONV = 3000
CLOWNS_ENONFINITE = 4001
CLOWNS_EBADNL = 4002
CLOWNS_EBADARG = 4003
CLOWNS_EHIP = 4004
CLOWNS_ENOMEM = 4005
CLOWNS_ESTRUCT = 4006
CLOWNS_EDECK = 4007
CLOWNS_EVOF = 4008

CLOWNS_MIXED = 1
CLOWNS_WARM_START = 2

ERRORS = {
    CLOWNS_NOCONV: "no convergence within ITER (reference IER=3000)",
    CLOWNS_ENONFINITE: "residual became NaN/Inf",
    CLOWNS_EBADNL: "NL must be 4 or 6",
    CLOWNS_EBADARG: "bad argument",
    CLOWNS_EHIP: "HIP runtime error",
    CLOWNS_ENOMEM: "device allocation failed",
    CLOWNS_ESTRUCT: "operator pattern is not structurally symmetric",
    CLOWNS_EDECK: "unreadable case file, or cells outside the box model",
    CLOWNS_EVOF: "VOF update: F above F_lim on a surface cell, or dt below DT_lim (the reference STOPs)",
}


class ClownsError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: clowns error {code} ({ERRORS.get(code, '?')})")


class Info(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int32),
        ("ier", ctypes.c_int32),
        ("refinements", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("err", ctypes.c_double),
        ("ms", ctypes.c_double),
        ("setup_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double

# name -> (restype, argtypes); mirrors include/clowns_pcg.h
SIGNATURES = {
    "clowns_pcgpme": (None, [_P] * 13),
    "clowns_solver_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, _P]),
    "clowns_solver_destroy": (ctypes.c_int, [_P]),
    "clowns_solver_set_stream": (ctypes.c_int, [_P, _P]),
    "clowns_solver_set_matrix": (ctypes.c_int, [_P, _P, _I, _I, _P, _P, _I, _D]),
    "clowns_solver_solve": (ctypes.c_int, [_P, _P, _P, _D, _I, _I, ctypes.POINTER(Info)]),
    "clowns_solver_solve_noalloc": (ctypes.c_int, [_P, _P, _P, _D, _I, _I, ctypes.POINTER(Info)]),
    "clowns_solver_get_scaled": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "clowns_solver_solve_transpose": (ctypes.c_int, [_P, _P, _P, _D, _I, _I, ctypes.POINTER(Info)]),
    "clowns_grad_coefficients": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "clowns_solver_time_kernel": (ctypes.c_int, [_P, _I, _I, ctypes.POINTER(_D)]),
    "clowns_solver_kernel_profile": (ctypes.c_int, [_P, _I, _P, _P]),
    "clowns_solver_patterns": (ctypes.c_int, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "clowns_pcgpme_will": (None, [_P] * 12),
    "clowns_solver_solve_will": (ctypes.c_int, [_P, _P, _P, _D, _I, ctypes.POINTER(Info), ctypes.POINTER(_I)]),
    "clowns_will_filter": (ctypes.c_int, [_P, _P, _I, _I, _I, _P, _P, _P, _P, ctypes.POINTER(_I)]),
    "clowns_solver_index_bits": (ctypes.c_int, [_P]),
    "clowns_solver_tuning": (ctypes.c_int, [_P, _P]),
    "clowns_solver_set_block_order": (ctypes.c_int, [_P, _P, _I]),
    "clowns_set_dx": (ctypes.c_int, [_P, _I, _P]),
    "clowns_deck_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _P]),
    "clowns_deck_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _P] + [_P] * 8),
    "clowns_deck_write_data": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _P, _P, _P, _D, _D, _I]),
    "clowns_deck_faces": (ctypes.c_int, [ctypes.c_char_p, _P, _P]),
    "clowns_deck_write_grid": (ctypes.c_int, [ctypes.c_char_p, _P, _P, _P, _I]),
    "clowns_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is not None and len(_lib._clowns_bound) != len(SIGNATURES):
        _bind(_lib)
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it (python -c 'import __graft_entry__ as g; g.build()' "
                "or make -C 2clowns-3d_amd); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        L._clowns_bound = set()
        _bind(L)
        _lib = L
    return _lib


def _bind(L):
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
        L._clowns_bound.add(name)


def check(rc, what):
    if rc != CLOWNS_OK:
        raise ClownsError(rc, what)
