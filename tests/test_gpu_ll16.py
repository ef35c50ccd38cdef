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
