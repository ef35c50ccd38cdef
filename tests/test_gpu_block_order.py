"""GPU: the order in which the FP64 SpMV and PM kernels visit their 512-row
blocks (clowns_solver_set_block_order, and the order the resident assembly
derives from the grid) changes nothing in the results: solves are
bit-identical to the natural order, the adjoint's transpose solve included.
Orders that are not a permutation of the kernels' blocks are refused."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _solver(block_order):
    from clowns3d import Solver
    os.environ["CLOWNS_BLOCK_ORDER"] = "1" if block_order else "0"
    try:
        return Solver(0)  # reads CLOWNS_BLOCK_ORDER at creation
    finally:
        del os.environ["CLOWNS_BLOCK_ORDER"]


@pytest.mark.parametrize("name", ["wavy32x24x16", "bathy28x24x16", "config1"])
def test_assembled_order_bit_identical(name):
    from clowns3d import DeviceGrid, configs
    from test_oracle_assembly import GRIDS
    # below 1M rows no order is built (a plane stays in L2 anyway): these
    # check that the switch changes nothing there; configurations 2 and 3
    # at full size, with the order, are in test_gpu_full_size.py
    g = configs.config1() if name == "config1" else GRIDS[name]()
    dg = DeviceGrid(g)
    out = []
    for on in (False, True):  # the last restores the default for later solvers
        sol = _solver(on)
        b = dg.assemble_into(sol, keep_raw=True)
        x64, i64 = sol.solve(b, eps=1e-10)
        xt, it = sol.solve(b, eps=2e-15, maxiter=3000)
        lam, il = sol.solve_transpose(b, eps=1e-10)  # the adjoint's solve, same order
        out.append((x64.cpu().numpy(), xt.cpu().numpy(), lam.cpu().numpy(), i64["iters"], it["iters"], il["iters"]))
        sol.close()
    assert out[0][3:] == out[1][3:]
    for a, c in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(a, c)


def test_set_block_order_bit_identical_and_checked():
    import torch
    from clowns3d import Solver
    from clowns3d._lib import CLOWNS_EBADARG, ClownsError
    s = box_system(160, 128, 6, seed=41)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    N = s["N"]
    G = ((N + 511) // 512 + 7) // 8 * 8
    sol = Solver(0)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x0, i0 = sol.solve(b, eps=1e-10)
    x0 = x0.cpu().numpy()
    perm = np.random.default_rng(3).permutation(G).astype(np.int32)
    sol.set_block_order(t(perm))
    x1, i1 = sol.solve(b, eps=1e-10)
    assert i1["iters"] == i0["iters"] and np.array_equal(x1.cpu().numpy(), x0)
    bad = perm.copy()
    bad[1] = bad[0]  # not a permutation
    for order in (bad, perm[:-8], np.append(perm, G).astype(np.int32)):
        with pytest.raises(ClownsError) as e:
            sol.set_block_order(t(order))
        assert e.value.code == CLOWNS_EBADARG
    sol.set_block_order(None)
    x2, _ = sol.solve(b, eps=1e-10)
    assert np.array_equal(x2.cpu().numpy(), x0)
    sol.close()

