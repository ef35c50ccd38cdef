"""GPU: clowns_solver_solve_noalloc, the solve for callers that cannot
allocate, against clowns_solver_solve on the same resident operator: the
same solution bit for bit and the same iterations, FP64 and mixed, on two
calls in a row with the second operator smaller; no change in the device's
free memory across a call; a row count past its maximum rejected."""
import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _pair(sol, b, mixed):
    import torch
    x_ref, i_ref = sol.solve(b, eps=1e-10, mixed=mixed)
    x = torch.zeros_like(b)
    x_na, i_na = sol.solve_noalloc(b, x, eps=1e-10, mixed=mixed)
    torch.cuda.synchronize()
    return x_ref, i_ref, x_na, i_na


def test_noalloc_matches_solve_on_two_calls():
    import torch
    from clowns3d import Solver
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol = Solver(0)
    for shape in [(64, 48, 16), (40, 30, 10)]:  # 49152 rows, then 12000
        s = box_system(*shape, seed=61)
        sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
        b = t(s["B"])
        for mixed in (False, True):
            x_ref, i_ref, x_na, i_na = _pair(sol, b, mixed)
            assert i_na["ier"] == i_ref["ier"] == 0
            assert i_na["iters"] == i_ref["iters"] and i_na["refinements"] == i_ref["refinements"]
            assert i_na["err"] == i_ref["err"]
            assert torch.equal(x_na, x_ref)
        # warm start from a perturbed solution
        x0 = x_ref * (1 + 1e-3)
        xw_ref, iw_ref = sol.solve(b, x0.clone(), eps=1e-10, warm=True)
        xw, iw = sol.solve_noalloc(b, x0.clone(), eps=1e-10, warm=True)
        assert iw["iters"] == iw_ref["iters"] and torch.equal(xw, xw_ref)
        # nothing allocated across a call (the array was set up by the first)
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(dev)[0]
        sol.solve_noalloc(b, x, eps=1e-10)
        sol.solve_noalloc(b, x, eps=1e-10, mixed=True)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info(dev)[0] == free0
    sol.close()


def test_noalloc_rejects_past_its_maximum():
    import torch
    from clowns3d import ClownsError, Solver
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    s = box_system(100, 90, 60, seed=62)  # 540000 rows > 524288
    sol = Solver(0)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    with pytest.raises(ClownsError) as e:
        sol.solve_noalloc(b, torch.zeros_like(b))
    assert e.value.code == 4003
    # mixed before the FP32 copies exist: rejected too
    s2 = box_system(20, 20, 10, seed=63)
    sol.set_matrix(t(s2["AL"].T), t(s2["LL"].T), t(s2["D"]))
    b2 = t(s2["B"])
    with pytest.raises(ClownsError):
        sol.solve_noalloc(b2, torch.zeros_like(b2), mixed=True)
    x, info = sol.solve_noalloc(b2, torch.zeros_like(b2))
    assert info["ier"] == 0
    sol.close()
