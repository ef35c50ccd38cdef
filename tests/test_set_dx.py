"""set_dx (src2LOWNS/1sub-main.F90:53-74) with the reference's round
(lib/src/xy_calc.f90:101-112): the product's host function clowns_set_dx
and the C oracle agree bit for bit with the flang build of the reference's
own set_dx (oracle/_ref), on uniform, non-uniform, large-coordinate and
tie cases and on the zero-last-face extrapolation; and the rounding rule
itself is pinned by hand-worked values (NINT rounds halves away from zero;
the number of decimals comes from the first cell size)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import REF_ASM_SO, oracle, ptr


def _oracle(x):
    x = np.array(x, dtype=np.float64)
    dx = np.zeros_like(x)
    assert oracle().oracle_set_dx(ptr(x), len(x), ptr(dx)) == 0
    return x, dx


def _ref(x):
    if not os.path.exists(REF_ASM_SO):
        pytest.skip("flang build of the reference (oracle/_ref) not present")
    lib = ctypes.CDLL(REF_ASM_SO, mode=os.RTLD_LAZY)
    x = np.array(x, dtype=np.float64)
    dx = np.zeros_like(x)
    lib.ref_set_dx(len(x), ptr(x), ptr(dx))
    return x, dx


def _product(x):
    from clowns3d import set_dx
    return set_dx(x)


def _cases():
    rng = np.random.default_rng(5)
    yield "uniform_0.1", np.arange(-2, 40) * 0.1
    yield "uniform_acc", np.cumsum(np.full(50, 0.1)) - 0.3
    yield "nonuniform", np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, 60))]) - 1.0
    yield "utm", 512345.0 + np.cumsum(np.round(rng.uniform(9.0, 11.0, 40), 3))
    yield "fine", np.cumsum(np.full(30, 0.0125)) + 3.0
    last = np.arange(-2, 20) * 0.25
    last[-1] = 0.0
    yield "zero_last", last
    yield "ties", np.array([0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0, 1.125])
    yield "tie_away", np.array([0.0, 2000.0, 4000.0, 6000.0, 8000.25, 10000.0, 12000.75, 14000.0])


@pytest.mark.parametrize("name,x", list(_cases()), ids=[c[0] for c in _cases()])
def test_set_dx_matches_reference(name, x):
    xo, do = _oracle(x)
    xp, dp = _product(x)
    assert np.array_equal(do, dp) and np.array_equal(xo, xp)
    xr, dr = _ref(x)
    assert np.array_equal(do, dr) and np.array_equal(xo, xr)


def test_round_rule_by_hand():
    # first cell x(ls+1)-x(ls) = 0.125: 0.125/10**(4-dec) > 1 first at
    # dec = 5, so a cell of 0.33333333 rounds to 0.33333
    x = [0.0, 0.125, 0.25, 0.375, 0.70833333, 1.0, 1.5]
    _, dx = _oracle(x)
    assert dx[2] == 0.125 and dx[3] == 0.33333
    # halves go away from zero: 2000 m cells give dec = 1, and a width of
    # 2000.25 (20002.5 tenths, exact) rounds to 2000.3, not to even 2000.2
    x = [0.0, 2000.0, 4000.0, 6000.0, 8000.25, 10000.0]
    _, dx = _oracle(x)
    assert dx[3] == 2000.3
    # the extrapolated last face
    x, dx = _oracle([0.0, 1.0, 2.0, 3.0, 4.0, 0.0])
    assert x[-1] == 5.0 and dx[-1] == dx[-2] == 1.0
