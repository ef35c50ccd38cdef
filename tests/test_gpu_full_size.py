"""GPU, configurations 2 and 3 at full size (256x256x64, 2.45M rows;
512x512x96 with bathymetry, 13.1M rows): the kernel forms the solver
chooses between per operator -- 8-bit against 16-bit neighbour codes (70k
and 155k wide slots), the grid's block order against the natural one --
give bit-identical FP64 solves on the real operators, and the solves meet
the reference's stopping rule."""
import os

import pytest

pytestmark = pytest.mark.gpu


def _solve(dg, env):
    from clowns3d import Solver
    for k, v in env.items():
        os.environ[k] = v
    try:
        sol = Solver(0)  # reads the switches at creation
    finally:
        for k in env:
            del os.environ[k]
    b = dg.assemble_into(sol)
    x, info = sol.solve(b, eps=1e-10)
    bits = sol.index_bits()
    sol.close()
    return x, info, bits


@pytest.mark.parametrize("cfg,rows", [(2, 2449992), (3, 13066133)])
def test_full_size_forms_bit_identical(cfg, rows):
    import torch
    from clowns3d import DeviceGrid, Solver, configs
    dg = DeviceGrid(configs.CONFIGS[cfg]())
    assert dg.N == rows
    x8, i8, b8 = _solve(dg, {"CLOWNS_LL8": "1", "CLOWNS_BLOCK_ORDER": "1"})
    x16, i16, b16 = _solve(dg, {"CLOWNS_LL8": "0", "CLOWNS_BLOCK_ORDER": "0"})
    # the defaults again for the solvers that follow
    for k, v in (("CLOWNS_LL8", "-1"), ("CLOWNS_BLOCK_ORDER", "1")):
        os.environ[k] = v
    try:
        Solver(0).close()
    finally:
        del os.environ["CLOWNS_LL8"], os.environ["CLOWNS_BLOCK_ORDER"]
    assert (b8, b16) == (8, 16)
    assert i8["ier"] == 0 and i8["err"] <= 1e-10 and i8["iters"] == i16["iters"]
    assert torch.equal(x8, x16)
