"""Classification (host, C++) and device assembly of the pressure system
through the C-ABI (include/clowns_pcg.h: clowns_classify, clowns_assemble,
clowns_solver_assemble).

``DeviceGrid(g)`` classifies a configs.Grid on the host (set_nflag, row
numbering, surface positions), uploads the fields to the GPU once and can then
assemble the PCGPME inputs there -- either in the reference layout
(``assemble``) or fused with the solver's scaling/sort straight into a
resident ``Solver`` (``assemble_into``).
"""
import ctypes

import numpy as np

from . import _lib as L


class GridDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("dt", ctypes.c_double)]


class GridDev(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("nrows", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("dt", ctypes.c_double), ("mp", ctypes.c_double)] + [
                   (n, ctypes.c_void_p) for n in ("nf", "nb", "mno", "row2cell", "sp_row", "F", "fb", "a", "un",
                                                  "dx", "xf")]


class PCInfo(ctypes.Structure):
    _fields_ = [("passes", ctypes.c_int32), ("iters", ctypes.c_int32), ("ier", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dmax_first", ctypes.c_double), ("dmax_last", ctypes.c_double),
                ("ms", ctypes.c_double)]


class VofInfo(ctypes.Structure):
    _fields_ = [("retries", ctypes.c_int32), ("modified", ctypes.c_int32), ("slid", ctypes.c_int32),
                ("rounds", ctypes.c_int32), ("dt", ctypes.c_double), ("ms", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int32
L.SIGNATURES.update({
    "clowns_classify": (ctypes.c_int, [ctypes.POINTER(GridDesc)] + [_P] * 11 + [_I, ctypes.POINTER(_I),
                                                                                 ctypes.POINTER(ctypes.c_double)]),
    "clowns_assemble": (ctypes.c_int, [ctypes.POINTER(GridDev), _I, _P, _P, _P, _P, _P]),
    "clowns_solver_assemble": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), ctypes.c_double, _P, _I]),
    "clowns_grad_open_ratios": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), _P, _P, _P, _P]),
    "clowns_solver_pressure_correction": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), _P, _P, ctypes.c_double, _I,
                                                         ctypes.c_double, _I, ctypes.POINTER(PCInfo)]),
    "clowns_solver_classify": (ctypes.c_int, [_P, ctypes.POINTER(GridDesc)] + [_P] * 11 +
                               [_I, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)]),
    "clowns_solver_vof": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), _P, _I, _P, _P, _P, _P,
                                         ctypes.POINTER(VofInfo)]),
    "clowns_solver_vof_time_kernel": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), _P, _I, _I, _P, _P,
                                                     ctypes.POINTER(ctypes.c_float)]),
    "clowns_solver_vof_plic": (ctypes.c_int, [_P, ctypes.POINTER(GridDev), _P, _I, _P, _P, _P, _P,
                                              ctypes.POINTER(VofInfo)]),
})


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def classify(g):
    """Host classification of a configs.Grid.  Returns a dict with nf, nb,
    mno (per cell), row2cell, sp_row (per row), N, mp and the (possibly
    clipped) F it classified."""
    A = g.arrays()
    nc = g.ncell
    out = dict(nf=np.zeros(nc, np.int8), nb=np.zeros(nc, np.int8), mno=np.zeros(nc, np.int32),
               row2cell=np.zeros(nc, np.int32), sp_row=np.zeros(nc), F=A["F"].copy())
    desc = GridDesc(g.nx, g.ny, g.nz, g.stride, g.dt)
    n = _I(0)
    mp = ctypes.c_double(0)
    L.check(L.lib().clowns_classify(ctypes.byref(desc), _p(A["solid"]), _p(out["F"]), _p(A["fb"]), _p(A["a"]),
                                    _p(A["dx"]), _p(A["xf"]), _p(out["nf"]), _p(out["nb"]), _p(out["mno"]),
                                    _p(out["row2cell"]), _p(out["sp_row"]), nc, ctypes.byref(n), ctypes.byref(mp)),
            "clowns_classify")
    out["N"] = n.value
    out["mp"] = mp.value
    out["row2cell"] = out["row2cell"][: n.value].copy()
    out["sp_row"] = out["sp_row"][: n.value].copy()
    out["arrays"] = A
    return out


def set_dx(x):
    """set_dx of one axis (src2LOWNS/1sub-main.F90:53-74) through
    clowns_set_dx: x holds the face coordinates x(ls-2 .. le+1).  Returns
    (x, dx) -- x with a zero last face extrapolated as the reference does."""
    x = np.array(x, dtype=np.float64)
    dx = np.zeros_like(x)
    L.check(L.lib().clowns_set_dx(_p(x), len(x), _p(dx)), "clowns_set_dx")
    return x, dx


class DeviceGrid:
    def __init__(self, g, device=0, classified=None):
        import torch
        self._torch = torch
        self.g = g
        c = classified if classified is not None else classify(g)
        self.N = c["N"]
        self.mp = c["mp"]
        self.host = c
        dev = torch.device(f"cuda:{device}")
        A = c["arrays"]
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.t = dict(nf=t(c["nf"]), nb=t(c["nb"]), mno=t(c["mno"]), row2cell=t(c["row2cell"]),
                      sp_row=t(c["sp_row"]), F=t(c["F"]), fb=t(A["fb"]), a=t(A["a"]), un=t(A["un"]),
                      dx=t(A["dx"]), xf=t(A["xf"]))
        ptrs = {k: ctypes.c_void_p(v.data_ptr()) for k, v in self.t.items()}
        self.dev = GridDev(g.nx, g.ny, g.nz, g.stride, self.N, 0, g.dt, self.mp, *[ptrs[k] for k in (
            "nf", "nb", "mno", "row2cell", "sp_row", "F", "fb", "a", "un", "dx", "xf")])
        self.device = dev

    @classmethod
    def on_device(cls, g, solver, device=0):
        """Upload the raw fields once and classify on the GPU
        (clowns_solver_classify); no host classification."""
        import torch
        self = cls.__new__(cls)
        self._torch = torch
        self.g = g
        dev = torch.device(f"cuda:{device}")
        self.device = dev
        A = g.arrays()
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        nc = g.ncell
        self.t = dict(solid=t(A["solid"]), F=t(A["F"]), fb=t(A["fb"]), a=t(A["a"]), un=t(A["un"]),
                      dx=t(A["dx"]), xf=t(A["xf"]),
                      nf=torch.empty(nc, dtype=torch.int8, device=dev),
                      nb=torch.empty(nc, dtype=torch.int8, device=dev),
                      mno=torch.empty(nc, dtype=torch.int32, device=dev),
                      row2cell=torch.empty(nc, dtype=torch.int32, device=dev),
                      sp_row=torch.empty(nc, dtype=torch.float64, device=dev))
        self.host = None
        self.N = 0
        self.mp = 0.0
        ptrs = {k: ctypes.c_void_p(v.data_ptr()) for k, v in self.t.items()}
        self.dev = GridDev(g.nx, g.ny, g.nz, g.stride, 0, 0, g.dt, 0.0, *[ptrs[k] for k in (
            "nf", "nb", "mno", "row2cell", "sp_row", "F", "fb", "a", "un", "dx", "xf")])
        self.classify_device(solver)
        return self

    def classify_device(self, solver):
        """(Re)classify on the GPU from the current device F.  Returns True
        if the exact host fallback was needed."""
        g = self.g
        T = self.t
        desc = GridDesc(g.nx, g.ny, g.nz, g.stride, g.dt)
        n = _I(0)
        mp = ctypes.c_double(0)
        fb = _I(0)
        p = lambda k: ctypes.c_void_p(T[k].data_ptr())
        L.check(L.lib().clowns_solver_classify(solver._h, ctypes.byref(desc), p("solid"), p("F"), p("fb"), p("a"),
                                               p("dx"), p("xf"), p("nf"), p("nb"), p("mno"), p("row2cell"),
                                               p("sp_row"), g.ncell, ctypes.byref(n), ctypes.byref(mp),
                                               ctypes.byref(fb)), "clowns_solver_classify")
        self.N = n.value
        self.mp = mp.value
        self.dev.nrows = self.N
        self.dev.mp = self.mp
        self.last_fallback = bool(fb.value)
        return self.last_fallback

    @property
    def row2cell(self):
        """Box index of each row (the first N entries of the row buffer)."""
        return self.t["row2cell"][: self.N]

    def update(self, g, classified):
        """New fill F (same static geometry: solid, fb, a, un, dx): upload F and
        the new classification (flags, numbering, surface positions)."""
        torch = self._torch
        c = classified
        self.g = g
        self.N = c["N"]
        self.host = c
        self.t["F"].copy_(torch.from_numpy(c["F"]))
        self.t["nf"].copy_(torch.from_numpy(c["nf"]))
        self.t["nb"].copy_(torch.from_numpy(c["nb"]))
        self.t["mno"].copy_(torch.from_numpy(c["mno"]))
        self.t["row2cell"] = torch.from_numpy(c["row2cell"]).to(self.device)
        self.t["sp_row"] = torch.from_numpy(c["sp_row"]).to(self.device)
        self.dev.nrows = self.N
        self.dev.row2cell = ctypes.c_void_p(self.t["row2cell"].data_ptr())
        self.dev.sp_row = ctypes.c_void_p(self.t["sp_row"].data_ptr())

    def set_velocity(self, un):
        """Replace the face velocities (e.g. between pressure-correction passes)."""
        self.t["un"].copy_(self._torch.as_tensor(un).reshape(self.t["un"].shape))

    def assemble(self, lda=None, stream=None):
        """Reference-layout PCGPME inputs: AL (6, lda), LL (6, lda) 1-based,
        D (N,), B (N,) as torch tensors on the device."""
        torch = self._torch
        lda = lda or self.N
        AL = torch.empty((6, lda), dtype=torch.float64, device=self.device)
        LL = torch.empty((6, lda), dtype=torch.int32, device=self.device)
        D = torch.empty(self.N, dtype=torch.float64, device=self.device)
        B = torch.empty(self.N, dtype=torch.float64, device=self.device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device).cuda_stream
        L.check(L.lib().clowns_assemble(ctypes.byref(self.dev), lda, ctypes.c_void_p(AL.data_ptr()),
                                        ctypes.c_void_p(LL.data_ptr()), ctypes.c_void_p(D.data_ptr()),
                                        ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(stream)),
                "clowns_assemble")
        return AL, LL, D, B

    def pressure_correction(self, solver, p=None, d0=1e-5, max_passes=1000, eps=1e-10, mixed=False):
        """The SMAC pressure-correction loop of one time step on the device
        (clowns_solver_pressure_correction): corrects the face velocities in
        self.t["un"] in place and the cell pressure p (zeros if None).
        Returns (p, info dict)."""
        torch = self._torch
        if p is None:
            p = torch.zeros(self.g.ncell, dtype=torch.float64, device=self.device)
        info = PCInfo()
        L.check(L.lib().clowns_solver_pressure_correction(
            solver._h, ctypes.byref(self.dev), ctypes.c_void_p(self.t["un"].data_ptr()),
            ctypes.c_void_p(p.data_ptr()), float(d0), int(max_passes), float(eps), 1 if mixed else 0,
            ctypes.byref(info)), "clowns_solver_pressure_correction")
        solver.N, solver.NL = self.N, 6
        solver.lda = (self.N + 63) // 64 * 64
        solver.version += 1
        return p, {k: getattr(info, k) for k, _ in PCInfo._fields_ if k != "reserved"}

    def vof(self, solver, Fp, mode=0, qs=None, qs_mask=None, dt=None, out_F=None, plic=None):
        """The VOF update cal_vof on the device (clowns_solver_vof), without
        its closing classification: face fluxes, surface stage (mode 1:
        quvwf_kaidan; mode 0: the PLIC face fluxes qs where qs_mask != 0,
        both (3, ncell), or none; plic = (cells, volumes (n, 6)): quvwf_normal
        from PLIC segment volumes, clowns_solver_vof_plic), quvwf_modify,
        cal_new_f, simple_slide(0).  Reads the flags, F, fb, a, un and dx held
        here; Fp is F of the previous step.  Returns (F_new, qf (3, ncell),
        info dict) with info["dt"] the step after cal_new_f's reductions."""
        torch = self._torch
        nc = self.g.ncell
        dev = self.device
        t = lambda a, dtype: (a if torch.is_tensor(a) else torch.from_numpy(np.ascontiguousarray(a))).to(
            dev, dtype).contiguous()
        Fp = t(Fp, torch.float64).view(-1)
        if Fp.numel() != nc:
            raise ValueError("Fp must hold one value per box cell")
        qf = torch.empty(3 * nc, dtype=torch.float64, device=dev)
        Fn = out_F if out_F is not None else torch.empty(nc, dtype=torch.float64, device=dev)
        qp = mp_ = None
        if qs is not None or qs_mask is not None:
            qs, qs_mask = t(qs, torch.float64).view(-1), t(qs_mask, torch.int8).view(-1)
            if qs.numel() != 3 * nc or qs_mask.numel() != 3 * nc:
                raise ValueError("qs and qs_mask must be (3, ncell)")
            qp, mp_ = ctypes.c_void_p(qs.data_ptr()), ctypes.c_void_p(qs_mask.data_ptr())
        g = GridDev.from_buffer_copy(self.dev)
        if dt is not None:
            g.dt = float(dt)
        info = VofInfo()
        if plic is not None:
            cells = t(plic[0], torch.int32).view(-1)
            vol = t(plic[1], torch.float64).view(-1)
            if vol.numel() != 6 * cells.numel():
                raise ValueError("plic volumes must be (n, 6)")
            L.check(L.lib().clowns_solver_vof_plic(solver._h, ctypes.byref(g), ctypes.c_void_p(Fp.data_ptr()),
                                                   int(cells.numel()), ctypes.c_void_p(cells.data_ptr()),
                                                   ctypes.c_void_p(vol.data_ptr()), ctypes.c_void_p(qf.data_ptr()),
                                                   ctypes.c_void_p(Fn.data_ptr()), ctypes.byref(info)),
                    "clowns_solver_vof_plic")
            return Fn, qf.view(3, nc), {k: getattr(info, k) for k, _ in VofInfo._fields_}
        L.check(L.lib().clowns_solver_vof(solver._h, ctypes.byref(g), ctypes.c_void_p(Fp.data_ptr()), int(mode),
                                          qp, mp_, ctypes.c_void_p(qf.data_ptr()), ctypes.c_void_p(Fn.data_ptr()),
                                          ctypes.byref(info)), "clowns_solver_vof")
        return Fn, qf.view(3, nc), {k: getattr(info, k) for k, _ in VofInfo._fields_}

    def vof_time_kernel(self, solver, Fp, kernel, reps=20, dt=None):
        """Average ms of one launch of a VOF-update kernel
        (clowns_solver_vof_time_kernel): 0 face fluxes, 1 new F, 2 slide,
        3 copy of F."""
        torch = self._torch
        nc = self.g.ncell
        Fp = (Fp if torch.is_tensor(Fp) else torch.from_numpy(np.ascontiguousarray(Fp))).to(
            self.device, torch.float64).contiguous()
        qf = torch.empty(3 * nc, dtype=torch.float64, device=self.device)
        Fn = torch.empty(nc, dtype=torch.float64, device=self.device)
        g = GridDev.from_buffer_copy(self.dev)
        if dt is not None:
            g.dt = float(dt)
        ms = ctypes.c_float(0.0)
        L.check(L.lib().clowns_solver_vof_time_kernel(solver._h, ctypes.byref(g), ctypes.c_void_p(Fp.data_ptr()),
                                                      int(kernel), int(reps), ctypes.c_void_p(qf.data_ptr()),
                                                      ctypes.c_void_p(Fn.data_ptr()), ctypes.byref(ms)),
                "clowns_solver_vof_time_kernel")
        return ms.value

    def with_open_ratios(self, a=None, fb=None):
        """A copy of this grid view whose face open ratios a (3, ncell) and
        cell open ratios fb (ncell,) are the given float64 CUDA tensors (kept
        alive by the view); the flags and numbering stay this grid's."""
        out = DeviceGrid.__new__(DeviceGrid)
        out.__dict__.update(self.__dict__)
        out.t = dict(self.t)
        fields = {k: getattr(self.dev, k) for k, _ in GridDev._fields_}
        for name, v in (("a", a), ("fb", fb)):
            if v is None:
                continue
            if v.dtype != self._torch.float64 or not v.is_cuda or not v.is_contiguous():
                raise TypeError(f"{name} must be a contiguous CUDA float64 tensor")
            if v.numel() != self.t[name].numel():
                raise ValueError(f"{name} has {v.numel()} entries, expected {self.t[name].numel()}")
            out.t[name] = v
            fields[name] = ctypes.c_void_p(v.data_ptr())
        out.dev = GridDev(**fields)
        return out

    def grad_open_ratios(self, solver, lam, x):
        """dL/da (3, ncell) and dL/dfb (ncell,) through this grid's assembly,
        from lambda = A^{-T} dL/dx and x (clowns_grad_open_ratios)."""
        torch = self._torch
        ga = torch.empty(self.t["a"].shape, dtype=torch.float64, device=self.device)
        gfb = torch.empty(self.t["fb"].shape, dtype=torch.float64, device=self.device)
        L.check(L.lib().clowns_grad_open_ratios(solver._h, ctypes.byref(self.dev), ctypes.c_void_p(lam.data_ptr()),
                                                ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(ga.data_ptr()),
                                                ctypes.c_void_p(gfb.data_ptr())), "clowns_grad_open_ratios")
        return ga, gfb

    def assemble_into(self, solver, parm=5.0, keep_raw=False, b=None):
        """Fused assembly + scaling + sort into the solver's resident operator.
        Returns the unscaled right-hand side b (N,)."""
        torch = self._torch
        if b is None:
            b = torch.empty(self.N, dtype=torch.float64, device=self.device)
        L.check(L.lib().clowns_solver_assemble(solver._h, ctypes.byref(self.dev), float(parm),
                                               ctypes.c_void_p(b.data_ptr()), 1 if keep_raw else 0),
                "clowns_solver_assemble")
        solver.N, solver.NL = self.N, 6
        solver.lda = (self.N + 63) // 64 * 64
        solver.version += 1
        return b
