"""GPU: the 16-bit coded neighbour indices of the two-row kernels
(k_ll16_build) gather exactly the elements the 32-bit LL does, so a solve
with them is bit-identical to one without them -- FP64 and mixed, on a box
whose z neighbours lie more than 16384 rows away (the far-offset code) and
with ghost slots."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _solve_all(s, ll16):
    import torch
    from clowns3d import Solver
    old = os.environ.get("CLOWNS_LL16")
    os.environ["CLOWNS_LL16"] = "1" if ll16 else "0"
    try:
        sol = Solver(0)  # reads CLOWNS_LL16 at creation
    finally:
        if old is None:
            del os.environ["CLOWNS_LL16"]
        else:
            os.environ["CLOWNS_LL16"] = old
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    xt, it = sol.solve_transpose(b, eps=1e-10)
    out = [x.cpu().numpy() for x in (x64, x32, xt)], (i64, i32, it)
    sol.close()
    return out


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9)])
def test_ll16_bit_identical(shape):
    s = box_system(*shape, seed=21)
    on, info_on = _solve_all(s, True)
    off, info_off = _solve_all(s, False)
    # restore the default for the tests that follow
    _solve_all(box_system(6, 5, 4, seed=1), True)
    for a, b in zip(on, off):
        assert np.array_equal(a, b)
    for a, b in zip(info_on, info_off):
        assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def test_ll16_wide_blocks_bit_identical():
    """Rows whose far neighbours spread more than the 16-bit code reaches
    within their 512-row block: that block stays on the 32-bit indices.
    Five iterations on a perturbed pattern, with and without the 16-bit
    form, must be bitwise equal."""
    import torch
    from clowns3d import Solver
    s = box_system(160, 128, 6, seed=22)
    N = s["N"]
    LL = s["LL"].copy(order="F")
    for i in range(3 * 512, 3 * 512 + 40):       # block 3: z+ 20480 -> 40480 away
        k = int(np.nonzero(LL[i] == i + 1 + 20480)[0][0])
        LL[i, k] = i + 1 + 40480
    for i in range(100 * 512, 100 * 512 + 40):   # block 100: z- -20480 -> -40480
        k = int(np.nonzero(LL[i] == i + 1 - 20480)[0][0])
        LL[i, k] = i + 1 - 40480
    assert LL[:N].min() >= 0 and LL[:N].max() <= N
    s["LL"] = LL
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    res, bits = [], []
    for ll16 in ("1", "0", "1"):
        os.environ["CLOWNS_LL16"] = ll16
        try:
            sol = Solver(0)
        finally:
            del os.environ["CLOWNS_LL16"]
        sol.set_matrix(t(s["AL"].T), t(LL.T), t(s["D"]))
        bits.append(sol.index_bits())
        x, info = sol.solve(t(s["B"]), eps=1e-30, maxiter=5)
        res.append(x.cpu().numpy())
        sol.close()
    assert bits[:2] == [16, 32]
    assert np.array_equal(res[0], res[1])
