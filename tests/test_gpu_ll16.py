"""GPU: the coded neighbour indices of the two-row kernels -- 16-bit
(k_ll16_build) and 8-bit (k_ll8_build) -- gather exactly the elements the
32-bit LL does, so a solve with either is bit-identical to one without them:
FP64 and mixed, on a box whose z neighbours lie more than 16384 rows away
(the far-offset code) and with ghost slots, and on a pattern whose offsets
spread past both codes' reach within a block (wide slots)."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu

# (CLOWNS_LL16, CLOWNS_LL8) -> index width the solver reports
FORMS = [(("1", "1"), 8), (("1", "0"), 16), (("0", "1"), 32)]
# the default: 8-bit codes built, their use timed per operator
DEFAULT = ("1", "-1")


def _solver(form):
    from clowns3d import Solver
    os.environ["CLOWNS_LL16"], os.environ["CLOWNS_LL8"] = form
    try:
        return Solver(0)  # reads the switches at creation
    finally:
        del os.environ["CLOWNS_LL16"]
        del os.environ["CLOWNS_LL8"]


def _restore_defaults():
    _solver(DEFAULT).close()


def _solve_all(s, form):
    import torch
    sol = _solver(form)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    xt, it = sol.solve_transpose(b, eps=1e-10)
    bits = sol.index_bits()
    tune = sol.tuning()
    # both precisions read the forced width (the FP32 row-major SpMV has
    # copies of the 8- and the 16-bit codes)
    assert tune["index64"] == bits and tune["index32"] == bits
    out = [x.cpu().numpy() for x in (x64, x32, xt)], (i64, i32, it), bits
    sol.close()
    return out


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9), (1100, 4, 4)])
def test_coded_indices_bit_identical(shape):
    s = box_system(*shape, seed=21)
    runs = [_solve_all(s, form) for form, _ in FORMS]
    _restore_defaults()
    assert [r[2] for r in runs] == [bits for _, bits in FORMS]
    for r in runs[1:]:
        for a, b in zip(runs[0][0], r[0]):
            assert np.array_equal(a, b)
        for a, b in zip(runs[0][1], r[1]):
            assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def test_wide_slots_bit_identical():
    """Rows whose far neighbours spread past the 16-bit code's reach (and so
    past the 8-bit code's 6-bit steps) within their 512-row block: those
    slots read the 32-bit LL.  Five iterations on a perturbed pattern in
    each index form must be bitwise equal."""
    import torch
    s = box_system(160, 128, 6, seed=22)
    N = s["N"]
    LL = s["LL"].copy(order="F")
    for i in range(3 * 512, 3 * 512 + 40):       # block 3: z+ 20480 -> 40480 away
        k = int(np.nonzero(LL[i] == i + 1 + 20480)[0][0])
        LL[i, k] = i + 1 + 40480
    for i in range(100 * 512, 100 * 512 + 40):   # block 100: z- -20480 -> -40480
        k = int(np.nonzero(LL[i] == i + 1 - 20480)[0][0])
        LL[i, k] = i + 1 - 40480
    for i in range(7 * 512, 7 * 512 + 8):        # block 7: y+ 160 -> 260 (past 6-bit steps)
        k = int(np.nonzero(LL[i] == i + 1 + 160)[0][0])
        LL[i, k] = i + 1 + 260
    assert LL[:N].min() >= 0 and LL[:N].max() <= N
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    res, bits = [], []
    for form, _ in FORMS:
        sol = _solver(form)
        sol.set_matrix(t(s["AL"].T), t(LL.T), t(s["D"]))
        x, info = sol.solve(t(s["B"]), eps=1e-30, maxiter=5)
        bits.append(sol.index_bits())
        res.append(x.cpu().numpy())
        sol.close()
    _restore_defaults()
    assert bits == [8, 16, 32]
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
