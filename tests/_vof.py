"""VOF-update test cases and the two checkers: the C oracle (oracle_vof) and
the flang build of the reference routines (oracle/_ref/libref_vof.so,
ref_vof).  The reference STOPs the program on some inputs (cal_new_f), so
it runs in a forked child process."""
import ctypes
import multiprocessing as mp
import os

import numpy as np

from conftest import ROOT, oracle, ptr

REF_VOF_SO = os.path.join(ROOT, "oracle", "_ref", "libref_vof.so")

VOF_OK, VOF_EVOF, VOF_ENONFINITE = 0, 4008, 4001


def vof_case(name, scenario="step", seed=0):
    """A grid with its classification inputs and a velocity field.

    scenario "step": velocities U(-0.1, 0.1) and the grid's dt;
    "stress": U(-1.5, 1.5), dt = 0.05 and F_prev = F perturbed, so that
    surface cells overfill (quvwf_modify) and cells slide (simple_slide);
    "retry": U(-3, 3), dt = 1: cal_new_f shrinks dt."""
    from test_oracle_assembly import GRIDS
    g = GRIDS[name]()
    rng = np.random.default_rng(seed)
    amp, dt = {"step": (0.1, g.dt), "stress": (1.5, 0.05), "retry": (3.0, 1.0)}[scenario]
    g.un[:] = rng.uniform(-amp, amp, size=g.un.shape)
    A = g.arrays()
    Fp = A["F"].copy()
    if scenario != "step":
        open_ = A["solid"] == 0
        Fp = np.where(open_, np.clip(Fp + rng.uniform(-0.05, 0.05, Fp.shape), 0.0, 1.0), 0.0)
    return g, A, Fp, dt


def plic_faces(g, A, nf, seed):
    """Stand-in PLIC face fluxes for NORMAL builds: on every face through
    which a surface cell drains (the faces quvwf_normal visits,
    1sub-vof.F90:138-151), a seeded value with the velocity's sign."""
    rng = np.random.default_rng(seed)
    nc = g.ncell
    st = (1, g.X, g.X * g.Y)
    un, a = A["un"], A["a"]
    qs = np.zeros((3, nc))
    mask = np.zeros((3, nc), np.int8)
    for c in np.nonzero(nf == 2)[0]:
        for l in range(3):
            for ll in (0, 1):
                n = c + ll * st[l]
                lsn = 1 - 2 * ll
                if un[l, n] * lsn >= 0 or a[l, n] == 0:
                    continue
                qs[l, n] = np.sign(un[l, n]) * rng.uniform(0.0, 1.2) * abs(un[l, n])
                mask[l, n] = 1
    return qs, mask


def plic_volumes(g, A, nf, seed, frac=1.0):
    """Stand-in PLIC segment volumes for NORMAL builds' quvwf_normal: for the
    listed surface cells (a seeded fraction of the nf = 2 cells) six
    volumes each, faces (l, ll) in the order 2 l + ll, up to the slab
    |u| dt * area; some zero.  Returns (cells int32, vol (n, 6))."""
    rng = np.random.default_rng(seed)
    cells = np.nonzero(nf == 2)[0]
    cells = cells[rng.random(len(cells)) < frac].astype(np.int32)
    vol = rng.uniform(0.0, 1.0, (len(cells), 6)) * 0.02
    vol[rng.random(vol.shape) < 0.1] = 0.0
    return cells, vol


def call_oracle_vof(g, A, nf, nb, F, Fp, dt, mode, qs=None, mask=None, plic=None):
    nc = g.ncell
    qf = np.zeros(3 * nc)
    Fn = np.zeros(nc)
    dtv = ctypes.c_double(dt)
    retries = ctypes.c_int(0)
    counts = np.zeros(2, np.int32)
    o = oracle()
    o.oracle_vof.restype = ctypes.c_int
    n_plic, pc, pv = 0, None, None
    if plic is not None:
        cells = np.ascontiguousarray(plic[0], dtype=np.int32)
        vol = np.ascontiguousarray(plic[1], dtype=np.float64)
        n_plic, pc, pv = len(cells), ptr(cells), ptr(vol)
    rc = o.oracle_vof(g.nx, g.ny, g.nz, g.stride, ptr(A["solid"]), ptr(nf), ptr(nb), ptr(F), ptr(Fp), ptr(A["fb"]),
                      ptr(A["a"]), ptr(A["un"]), ptr(A["dx"]), ctypes.byref(dtv), mode,
                      ptr(qs) if qs is not None else None, ptr(mask) if mask is not None else None, n_plic, pc, pv,
                      ptr(qf), ptr(Fn), ctypes.byref(retries), ptr(counts))
    return dict(rc=rc, qf=qf.reshape(3, nc), F=Fn, dt=dtv.value, retries=retries.value, modified=int(counts[0]),
                slid=int(counts[1]))


def _ref_child(args):
    nx, ny, nz, stride, A, F, Fp, dt, mode, qs, mask = args
    lib = ctypes.CDLL(REF_VOF_SO, mode=os.RTLD_LAZY)
    nc = (nx + 2) * (ny + 2) * (nz + 2)
    nf = np.zeros(nc, np.int8)
    nb = np.zeros(nc, np.int8)
    fcls = np.zeros(nc)
    qf = np.zeros(3 * nc)
    fo = np.zeros(nc)
    dtv = ctypes.c_double(dt)
    secs = ctypes.c_double(0.0)
    qs = np.zeros(3 * nc) if qs is None else np.ascontiguousarray(qs).ravel()
    mask = np.zeros(3 * nc, np.int8) if mask is None else np.ascontiguousarray(mask).ravel()
    lib.ref_vof(nx, ny, nz, ptr(A["solid"]), ptr(F), ptr(Fp), ptr(A["fb"]), ptr(A["a"]), ptr(A["un"]),
                ptr(A["dx"]), ptr(A["xf"]), stride, ctypes.byref(dtv), mode, ptr(qs), ptr(mask), ptr(nf), ptr(nb),
                ptr(fcls), ptr(qf), ptr(fo), 1, ctypes.byref(secs))
    return dict(nf=nf, nb=nb, F_cls=fcls, qf=qf.reshape(3, nc), F=fo, dt=dtv.value)


def call_ref_vof(g, A, F, Fp, dt, mode, qs=None, mask=None):
    """ref_vof in a forked child (a STOP there ends only the child).
    Returns None when the child did not come back."""
    ctx = mp.get_context("fork")
    with ctx.Pool(1) as pool:
        res = pool.apply_async(_ref_child, ((g.nx, g.ny, g.nz, g.stride, A, F, Fp, dt, mode, qs, mask),))
        try:
            return res.get(timeout=120)
        except Exception:
            return None
