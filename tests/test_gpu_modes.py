"""GPU: mixed-precision solve and the adjoint (transpose) solve.

Mixed: FP32 inner PM-Bi-CGSTAB + FP64 refinement must reach the reference's
stopping value (sqrt(RN/BN) <= 1e-10, here on the true residual) and be as
accurate as the FP64 solve: error against a tight FP64 solve within
max(1e-9, 2x the FP64 solve's), with the inner preconditioner's
coefficients in half precision and in FP32.  Adjoint: lambda = A^-T g must
satisfy A^T lambda = g (checked with a host sparse product of the unscaled
transpose) and the duality <g, A^-1 b> = <A^-T g, b>.
"""
import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _resident(s, pm_half=None):
    import os
    import torch
    from clowns3d import Solver
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if pm_half is None:
        sol = Solver(0)
    else:
        os.environ["CLOWNS_PM_HALF"] = "1" if pm_half else "0"
        try:
            sol = Solver(0)  # reads CLOWNS_PM_HALF at creation
        finally:
            del os.environ["CLOWNS_PM_HALF"]
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    return sol, t


def apply_AT(s, y):
    N, NL = s["N"], s["NL"]
    out = s["D"] * y
    for k in range(NL):
        col = s["LL"][:N, k]
        m = col > 0
        np.add.at(out, col[m] - 1, s["AL"][:N, k][m] * y[:N][m])
    return out


# pm_half: the inner preconditioner's coefficients in half precision (the
# default) or FP32; the last case restores the default for later solvers
@pytest.mark.parametrize("pm_half", [False, True])
@pytest.mark.parametrize("shape", [(33, 17, 9), (64, 48, 16)])
def test_mixed_matches_fp64(shape, pm_half):
    s = box_system(*shape, seed=11)
    sol, t = _resident(s, pm_half)
    b = t(s["B"])
    x64, i64 = sol.solve(b)
    x32, i32 = sol.solve(b, mixed=True)
    xt, it = sol.solve(b, eps=2e-15)  # tight FP64 solve (stagnation exit)
    assert i64["ier"] == 0 and i32["ier"] == 0, (i64, i32)
    assert i32["err"] <= 1e-10
    err = lambda x: (x - xt).norm().item() / xt.norm().item()
    print("error vs tight: mixed", err(x32), "fp64", err(x64), i32, i64)
    # both stop at the same rule; mixed must be as accurate as FP64
    assert err(x32) <= max(1e-9, 2.0 * err(x64))
    sol.close()


def test_transpose_solve():
    s = box_system(33, 17, 9, seed=12)
    sol, t = _resident(s)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(s["N"])
    lam, info = sol.solve_transpose(t(g), eps=1e-12)
    assert info["ier"] == 0, info
    lam = lam.cpu().numpy()
    r = apply_AT(s, lam) - g
    assert np.linalg.norm(r) / np.linalg.norm(g) < 1e-8
    x, _ = sol.solve(t(s["B"]), eps=1e-12)
    x = x.cpu().numpy()
    lhs, rhs = g @ x, lam @ s["B"]
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs)
    sol.close()
