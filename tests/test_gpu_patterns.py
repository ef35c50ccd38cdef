"""GPU: row patterns.  The SpMV of both precisions reads a one-byte pattern
id for rows that repeat a table row instead of the coefficients; the results must be bit-identical to the operator without
patterns -- FP64, mixed and transposed solves -- on an assembled operator
(most rows interior and uniform) and on the drop-in's copy of it, and a
table kept from one operator must give correct ids on the next."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(on):
    from clowns3d import Solver
    os.environ["CLOWNS_PATTERNS"] = "1" if on else "0"
    try:
        return Solver(0)  # reads the switch at creation
    finally:
        del os.environ["CLOWNS_PATTERNS"]


def _grid(shift=0.0, bathymetry=False):
    from clowns3d import configs
    if bathymetry:
        g = configs.config_bathymetry(192, 160, 40, nbumps=12)
    else:
        g = configs.config_wavy_box(160, 128, 40, (60, 100, 50, 80, 0, 14))
    if shift:
        g.set_water(configs.wavy_surface(g) + shift)
    return g


def _solve(dg, on):
    import torch
    sol = _solver(on)
    b = dg.assemble_into(sol, keep_raw=True)
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    xt, it = sol.solve_transpose(b, eps=1e-10)
    pats = sol.patterns()
    torch.cuda.synchronize()
    out = [x.cpu().numpy() for x in (x64, x32, xt)], (i64, i32, it), pats
    sol.close()
    return out


@pytest.mark.parametrize("bathymetry", [False, True])
def test_patterns_bit_identical(bathymetry):
    from clowns3d import DeviceGrid
    dg = DeviceGrid(_grid(bathymetry=bathymetry))
    on, info_on, pats = _solve(dg, True)
    off, info_off, pats_off = _solve(dg, False)
    print(f"N={dg.N}: {pats[0]} patterns cover {pats[1]} rows ({pats[1] / dg.N:.3f})")
    assert pats_off == (0, 0)
    assert pats[0] > 0 and pats[1] > dg.N // 2
    for a, b in zip(on, off):
        assert np.array_equal(a, b)
    for a, b in zip(info_on, info_off):
        assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def test_table_kept_for_the_next_operator():
    """A second operator (the surface moved) takes its ids from the first
    operator's table; solves equal those of a fresh solver without patterns."""
    import torch
    from clowns3d import DeviceGrid
    sol = _solver(True)
    dg1 = DeviceGrid(_grid())
    b1 = dg1.assemble_into(sol)
    sol.solve(b1, eps=1e-10)
    dg2 = DeviceGrid(_grid(shift=0.37))
    b2 = dg2.assemble_into(sol)
    x, info = sol.solve(b2, eps=1e-10)
    xm, im = sol.solve(b2, eps=1e-10, mixed=True)
    pats = sol.patterns()
    ref = _solver(False)
    b2r = dg2.assemble_into(ref)
    xr, ir = ref.solve(b2r, eps=1e-10)
    xmr, imr = ref.solve(b2r, eps=1e-10, mixed=True)
    torch.cuda.synchronize()
    sol.close()
    ref.close()
    assert pats[1] > dg2.N // 2
    assert info["iters"] == ir["iters"] and im["iters"] == imr["iters"]
    assert torch.equal(x, xr) and torch.equal(xm, xmr)


def _solve_env(dg, env):
    import torch
    from clowns3d import Solver
    os.environ.update(env)
    try:
        sol = Solver(0)  # reads the switches at creation
    finally:
        for k in env:
            del os.environ[k]
    b = dg.assemble_into(sol, keep_raw=True)
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    tune = sol.tuning()
    torch.cuda.synchronize()
    out = (x64.cpu().numpy(), x32.cpu().numpy()), (i64, i32), tune
    sol.close()
    return out


@pytest.mark.parametrize("env", [
    {"CLOWNS_STREAM": "1", "CLOWNS_LL8": "1", "CLOWNS_FUSE_PM": "0"},
    {"CLOWNS_STREAM": "0", "CLOWNS_LL8": "0", "CLOWNS_FUSE_PM": "0"},
    {"CLOWNS_STREAM": "2", "CLOWNS_LL8": "1", "CLOWNS_FUSE_PM": "1"},
    {"CLOWNS_STREAM": "0", "CLOWNS_LL8": "1", "CLOWNS_FUSE_PM": "0", "CLOWNS_BLOCK_ORDER": "0"},
], ids=["streamed-8bit", "cached-16bit", "pairs-8bit", "natural-order"])
def test_pattern_pm_stages_bit_identical(env):
    """The FP64 PM stages with the pattern table, two 128-row chunks a
    wavefront (k_pm_p2: single stages and the two-term stage with its dots),
    compute what k_pm_v2 does: FP64 and mixed solves bit-identical with the
    form forced on and off, in the grid's block order and in the natural one,
    with 8- and 16-bit codes, streamed and cached; the solver reports it."""
    from clowns3d import DeviceGrid
    dg = DeviceGrid(_grid(bathymetry=True))
    on = _solve_env(dg, dict(env, CLOWNS_PM_PAT="1"))
    off = _solve_env(dg, dict(env, CLOWNS_PM_PAT="0"))
    print(f"N={dg.N}: {on[2]} | {off[2]}")
    assert on[2]["pmpat64"] == 1 and off[2]["pmpat64"] == 0
    for a, b in zip(on[0], off[0]):
        assert np.array_equal(a, b)
    for a, b in zip(on[1], off[1]):
        assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def test_pattern_pm_stages_timed_choice():
    """By default the solver times the form with the others (tune_stream) and
    reports its choice; the solve equals the forced forms'."""
    from clowns3d import DeviceGrid
    dg = DeviceGrid(_grid())
    tm = _solve_env(dg, {})
    ref = _solve_env(dg, {"CLOWNS_PM_PAT": "0"})
    print(f"N={dg.N}: timed {tm[2]}")
    assert tm[2]["pmpat64"] in (0, 1)
    for a, b in zip(tm[0], ref[0]):
        assert np.array_equal(a, b)
