"""GPU: the non-temporal (streamed) forms of the PM stages and the SpMV, and
the fused PM stage pairs, compute exactly what the cached and single forms
do -- FP64, mixed (and transposed) solves bit-identical with the forms
forced on and off -- and the solver reports the forms it chose."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def _solver(stream):
    from clowns3d import Solver
    os.environ["CLOWNS_STREAM"] = str(stream)
    try:
        return Solver(0)  # reads CLOWNS_STREAM at creation
    finally:
        del os.environ["CLOWNS_STREAM"]


def _solve_all(s, stream):
    import torch
    sol = _solver(stream)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    xt, it = sol.solve_transpose(b, eps=1e-10)
    tune = sol.tuning()
    out = [x.cpu().numpy() for x in (x64, x32, xt)], (i64, i32, it), tune
    sol.close()
    return out


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9)])
def test_streamed_forms_bit_identical(shape):
    """CLOWNS_STREAM: 1 both streamed, 0 neither, 2 the PM stages only, 3
    the SpMV only (the solver reports bit 0 PM, bit 1 SpMV), -1 timed."""
    s = box_system(*shape, seed=23)
    runs = {st: _solve_all(s, st) for st in (1, 0, 2, 3, -1)}
    for st, bits in ((1, 3), (0, 0), (2, 1), (3, 2)):
        assert runs[st][2]["stream64"] == bits and runs[st][2]["stream32"] == bits
    assert runs[-1][2]["stream64"] in (0, 1, 2, 3) and runs[-1][2]["fuse64"] in (0, 1)
    ref, ref_info, _ = runs[1]
    for st in (0, 2, 3, -1):
        for a, b in zip(ref, runs[st][0]):
            assert np.array_equal(a, b)
        for a, b in zip(ref_info, runs[st][1]):
            assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def _solve_fused(s, fuse):
    import torch
    from clowns3d import Solver
    os.environ["CLOWNS_FUSE_PM"] = str(fuse)
    try:
        sol = Solver(0)  # reads CLOWNS_FUSE_PM at creation
    finally:
        del os.environ["CLOWNS_FUSE_PM"]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    b = t(s["B"])
    x64, i64 = sol.solve(b, eps=1e-10)
    x32, i32 = sol.solve(b, eps=1e-10, mixed=True)
    tune = sol.tuning()
    sol.close()
    return [x64.cpu().numpy(), x32.cpu().numpy()], (i64, i32), tune


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9)])
def test_fused_pm_pairs_bit_identical(shape):
    """Two PM stages fused in one launch (k_pm2_v2 recomputes the first stage
    at the second's gather point) compute what the two single stages do."""
    s = box_system(*shape, seed=24)
    on, info_on, tune_on = _solve_fused(s, 1)
    off, info_off, tune_off = _solve_fused(s, 0)
    _solve_fused(box_system(6, 5, 4, seed=1), -1)  # back to timing per operator
    assert tune_on["fuse64"] == 1 and tune_off["fuse64"] == 0
    for a, b in zip(on, off):
        assert np.array_equal(a, b)
    for a, b in zip(info_on, info_off):
        assert a["iters"] == b["iters"] and a["ier"] == b["ier"] == 0


def _solve_v4(s, v4):
    import torch
    from clowns3d import Solver
    os.environ["CLOWNS_V4_32"] = str(v4)
    os.environ["CLOWNS_FUSE_PM"] = "0"  # single stages, the ones with four rows a thread
    try:
        sol = Solver(0)  # reads both at creation
    finally:
        del os.environ["CLOWNS_V4_32"]
        os.environ["CLOWNS_FUSE_PM"] = "-1"
        Solver(0).close()  # the pair/single choice timed again from here on
        del os.environ["CLOWNS_FUSE_PM"]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    x32, i32 = sol.solve(t(s["B"]), eps=1e-10, mixed=True)
    sol.close()
    return x32.cpu().numpy(), i32


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9)])
def test_four_row_fp32_stages_bit_identical(shape):
    """The FP32 single PM stages with four rows a thread (k_pm_v4, 5049 rows:
    a ragged tail) compute what two rows a thread do."""
    s = box_system(*shape, seed=25)
    a, ia = _solve_v4(s, 1)
    b, ib = _solve_v4(s, 0)
    _solve_v4(box_system(6, 5, 4, seed=1), 1)  # the default again
    assert ia["iters"] == ib["iters"] and ia["ier"] == ib["ier"] == 0
    assert np.array_equal(a, b)


def _solve_rows(s, rows):
    import torch
    from clowns3d import Solver
    os.environ["CLOWNS_SPMV_ROWS"] = str(rows)
    try:
        sol = Solver(0)  # reads CLOWNS_SPMV_ROWS at creation
    finally:
        del os.environ["CLOWNS_SPMV_ROWS"]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    x32, i32 = sol.solve(t(s["B"]), eps=1e-10, mixed=True)
    sol.close()
    return x32.cpu().numpy(), i32


@pytest.mark.parametrize("shape", [(160, 128, 6), (33, 17, 9)])
def test_row_major_fp32_spmv_bit_identical(shape):
    """The FP32 SpMV from row-major copies of the coefficients and 16-bit
    codes (k_spmv_rows32; 5049 rows: an odd tail row) computes what the
    column-major one does."""
    s = box_system(*shape, seed=26)
    a, ia = _solve_rows(s, 1)
    b, ib = _solve_rows(s, 0)
    _solve_rows(box_system(6, 5, 4, seed=1), 1)  # the default again
    assert ia["iters"] == ib["iters"] and ia["ier"] == ib["ier"] == 0
    assert np.array_equal(a, b)
