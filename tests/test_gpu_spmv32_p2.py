"""GPU: the FP32 SpMV of the mixed inner solve in the two-chunk form
(k_spmv_rows32_p2: both of a wavefront's 128-row chunks of ids, codes and x
issued before the first chunk's gathers) computes what the one-chunk form
(k_spmv_rows32) does.  Mixed solves are bit-identical with the form switched
on and off, with 8- and 16-bit codes: on an assembled operator with and
without the row-pattern table, cached and streamed; on a ragged system with
an odd tail row; and on one whose scaled FP32 diagonal is not 1 on some rows
(the diagonal stream is read)."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _solver(env):
    from clowns3d import Solver
    os.environ.update(env)
    try:
        return Solver(0)  # reads the switches at creation
    finally:
        for k in env:
            del os.environ[k]


def _mixed(env, dg=None, s=None):
    import torch
    sol = _solver(env)
    if dg is not None:
        b = dg.assemble_into(sol, keep_raw=True)
    else:
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")
        sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
        b = t(s["B"])
    x, info = sol.solve(b, eps=1e-10, mixed=True)
    tune = sol.tuning()
    torch.cuda.synchronize()
    out = x.cpu().numpy(), info, tune
    sol.close()
    return out


def _both(env, converge=True, **kw):
    on = _mixed(dict(env, CLOWNS_SPMV32_P2="1"), **kw)
    off = _mixed(dict(env, CLOWNS_SPMV32_P2="0"), **kw)
    print(f"  two-chunk {on[2]} | one-chunk {off[2]}; {on[1]['iters']} iterations", flush=True)
    assert on[2]["index32"] == (8 if env["CLOWNS_LL8"] == "1" else 16) and (on[2]["spmv32"] & 3) == 2
    assert (off[2]["spmv32"] & 3) == 1
    assert on[1]["iters"] == off[1]["iters"] and on[1]["ier"] == off[1]["ier"]
    assert on[1]["ier"] == 0 or not converge
    assert np.array_equal(on[0], off[0])
    return on


@pytest.mark.parametrize("ll8", ["1", "0"], ids=["8bit", "16bit"])
@pytest.mark.parametrize("patterns", ["1", "0"])
@pytest.mark.parametrize("stream", ["0", "1"], ids=["cached", "streamed"])
def test_assembled_operator(patterns, stream, ll8):
    from clowns3d import DeviceGrid, configs
    dg = DeviceGrid(configs.config_wavy_box(160, 128, 40, (60, 100, 50, 80, 0, 14)))
    print(f"N={dg.N}", flush=True)
    on = _both({"CLOWNS_LL8": ll8, "CLOWNS_STREAM": stream, "CLOWNS_PATTERNS": patterns}, dg=dg)
    assert bool(on[2]["spmv32"] & 4) == (patterns == "1")


@pytest.mark.parametrize("ll8", ["1", "0"], ids=["8bit", "16bit"])
def test_odd_tail_row(ll8):
    s = box_system(33, 17, 9, seed=27)
    assert s["N"] % 2 == 1
    _both({"CLOWNS_LL8": ll8, "CLOWNS_STREAM": "0"}, s=s)


@pytest.mark.parametrize("ll8", ["1", "0"], ids=["8bit", "16bit"])
def test_non_unit_diagonal(ll8):
    s = box_system(64, 48, 10, seed=28)
    N = s["N"]
    off = np.abs(s["AL"][:N]).sum(axis=1)
    rows = np.arange(5, N, 97)
    s["D"][rows] = 0.9 * off[rows]  # max(sum|AL|, D) = sum|AL|: D W != 1
    _both({"CLOWNS_LL8": ll8, "CLOWNS_STREAM": "1"}, converge=False, s=s)
