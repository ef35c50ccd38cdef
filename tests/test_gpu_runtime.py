"""GPU: the solver's runtime around the kernels.

* Ghost rows: the drop-in keeps one solver per device (src2LOWNS calls
  PCGPME with a different mnoe every pass and step).  A large solve that
  ends non-finite leaves NaN in its work vectors past a smaller N; the
  smaller solve that follows must still read zeros at its ghost rows
  N+1+mod(I,64) (pcgpme_para.F90:80), also where the caller's AL is not 0
  in a slot with LL = 0, and match the oracle.
* In-solve kernel profile: the launch counts per kernel class follow the
  iteration (pcgpme_para.F90:148-247) exactly, and a profiled solve gives
  the same solution as an unprofiled one.
"""
import os

import numpy as np
import pytest

from conftest import call_oracle, copy_system
from _systems import box_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def drop_in(s, iters=10000, eps=1e-10):
    import clowns3d
    c = copy_system(s)
    N = c["N"]
    X = np.zeros(N + 64)
    EPS = np.array([eps, 0.0])
    it, ier = clowns3d.pcgpme(c["AL"], 6, c["NA"], c["LL"], c["D"], N, 64, c["B"], X, EPS, 5.0, iters)
    return X[:N], it, ier, c


def test_ghost_rows_after_a_larger_nonfinite_solve():
    big = box_system(40, 30, 20, seed=41)
    big["B"] = big["B"] * 1e300  # BN overflows: the solve ends non-finite
    _, _, ier, _ = drop_in(big)
    assert ier == 4001
    small = box_system(24, 20, 10, seed=42)
    N = small["N"]
    # a caller's nonzero coefficient in every slot without a neighbour
    for k in range(6):
        m = small["LL"][:N, k] == 0
        small["AL"][:N, k][m] = -0.01
    for iters in (1, 2, 5):
        x, it, ier, c = drop_in(small, iters=iters)
        o = call_oracle(small, iters=iters)
        assert ier == o["IER"] == 3000 and it == o["ITER"]
        assert np.isfinite(x).all()
        assert rel(x, o["X"][:N]) < 1e-12
        assert rel(c["B"], o["B"]) < 1e-12
    x, it, ier, _ = drop_in(small)
    o = call_oracle(small)
    assert ier == o["IER"] == 0 and abs(it - o["ITER"]) <= 2
    assert rel(x, o["X"][:N]) < 1e-8


def test_kernel_profile_counts():
    import torch
    from clowns3d import Solver
    os.environ["CLOWNS_FUSE_PM"] = "0"  # single PM stages: the counts below
    try:
        sol = Solver(0)
    finally:
        del os.environ["CLOWNS_FUSE_PM"]
    s = box_system(64, 48, 16, seed=43)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x0, i0 = sol.solve(b, eps=1e-10)
    sol.kernel_profile(True)
    x1, i1 = sol.solve(b, eps=1e-10)
    p = sol.kernel_profile(False)
    sol.kernel_profile(True)
    xm, im = sol.solve(b, eps=1e-10, mixed=True)
    q = sol.kernel_profile(False)
    sol.close()
    assert torch.equal(x0, x1) and i0["iters"] == i1["iters"]
    K = i1["iters"]
    f = {c: v[0] for c, v in p["fp64"].items()}
    # setup: one PM application of W b (four single stages and stage 5/6)
    assert f["spmv"] == 2 * K and f["residual_spmv"] == 0
    assert f["pm_stage"] == 8 * K + 4 and f["pm_dot"] == 2 * K + 1 and f["pm_pair"] == 0
    assert f["p_update"] == K - 1 and f["e_update"] == K and f["xr_update"] == K
    assert f["finalize"] == 3 * K
    assert all(v == (0, 0.0) for v in p["fp32"].values())
    assert all(v[1] > 0 for v in p["fp64"].values() if v[0] > 0)
    # mixed: the FP64 side is the PM application of the right-hand side and,
    # after every refinement, the true residual and its PM application
    R = im["refinements"]
    f = {c: v[0] for c, v in q["fp64"].items()}
    assert f["residual_spmv"] == R and f["spmv"] == 0
    assert f["pm_stage"] == 4 * (R + 1) and f["pm_dot"] == R + 1
    g = {c: v[0] for c, v in q["fp32"].items()}
    Ki = im["iters"]
    assert g["spmv"] == 2 * Ki and g["pm_stage"] == 8 * Ki and g["pm_dot"] == 2 * Ki
    assert g["xr_update"] == Ki and g["finalize"] == 3 * Ki
