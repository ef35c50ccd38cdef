"""Pins the C oracle of PCGPME_WILL (oracle_pcgpme_will, oracle_will_filter in
oracle/oracle_pcgpme.c), the 2DH dispersion solve of
src2LOWNS/pcgpme_2DH.f90.

1. Golden fixtures written by the flang build of the reference's own
   pcgpme_2DH.f90 (tests/golden/make_golden_will.py): B on return bit for bit
   at one thread, for ITER = 0, 1, 2, 5 and to convergence.
2. The flang build itself (when oracle/_ref is present) on fresh 2DH
   systems, and the inputs AL, LL, D left untouched by both.
3. The output pass on hand-worked cases whose serial answer is fixed by the
   rules of pcgpme_2DH.f90:255-268 (zero limit, snap to a neighbour in slot
   order reading already filtered lower rows, self slot skipped, first zero
   slot ends the row, B keeps the residual where no neighbour differs), and
   on the realistic systems: every branch is taken, including chains where a
   row snaps to a lower row that itself snapped.
"""
import ctypes
import glob
import os

import numpy as np
import pytest

from conftest import ROOT, call_oracle_will, call_ref_will, oracle, ptr

GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "pcgpme_will_*.npz")))


def _from_golden(z):
    return dict(AL=np.asfortranarray(z["AL_in"]), LL=np.asfortranarray(z["LL_in"]), D=z["D_in"].copy(),
                B=z["B_in"].copy(), N=int(z["N"]), NA=int(z["NA"]), NL=int(z["NL"]))


def test_golden_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
@pytest.mark.parametrize("iters", [0, 1, 2, 5, 1000])
def test_oracle_matches_reference_golden(path, iters):
    z = np.load(path)
    s = _from_golden(z)
    o = call_oracle_will(s, iters=iters)
    assert o["rc"] == 0
    assert np.array_equal(o["B"], z[f"it{iters}_B"])
    for k in ("AL", "LL", "D"):
        assert np.array_equal(o[k], s[k]), k  # inputs only


@pytest.mark.parametrize("shape,seed", [((64, 48), 41), ((130, 90), 42)])
def test_oracle_matches_flang_reference(shape, seed):
    from clowns3d import configs
    s = configs.dispersion_2dh(*shape, seed=seed)
    for iters in (0, 3, 1000):
        o = call_oracle_will(s, iters=iters)
        r = call_ref_will(s, iters=iters)
        assert np.array_equal(o["B"], r["B"]), iters
        for k in ("AL", "LL", "D"):
            assert np.array_equal(r[k], s[k]) and np.array_equal(o[k], s[k]), k


def _filter(LL, X, B):
    LL = np.asfortranarray(np.asarray(LL, dtype=np.int32))
    X = np.array(X, dtype=np.float64)
    B = np.array(B, dtype=np.float64)
    N, NL = LL.shape
    oracle().oracle_will_filter(ptr(LL), NL, N, N, ptr(X), ptr(B))
    return X, B


def test_filter_hand_cases():
    e = 1e-14
    # 1) a chain along 1-2-3: row 1 snaps to row 2's raw value, row 2 then
    #    equals row 1 (already filtered) and snaps to row 3, row 3 equals row 2
    x = [1e-13, 1e-13 + 0.5 * e, 1e-13 + 0.9 * e]
    LL = [[2, 0], [1, 3], [2, 0]]
    X, B = _filter(LL, x, [7.0, 8.0, 9.0])
    assert X[0] == x[1] and X[1] == x[2] and X[2] == x[2]
    # B(i) is stored only past a differing neighbour: row 3's only
    # neighbour equals it, so B(3) keeps the residual
    assert B[0] == x[1] and B[1] == x[2] and B[2] == 9.0
    # 2) the zero limit, then a snap back to the raw (not yet zeroed) tiny
    #    value of a higher row; an exact-zero neighbourhood keeps the residual
    X, B = _filter([[2], [1]], [0.9 * e, -0.5 * e], [3.0, 4.0])
    assert X[0] == -0.5 * e and B[0] == -0.5 * e
    assert X[1] == -0.5 * e and B[1] == -0.5 * e  # zeroed, then snapped to row 1
    X, B = _filter([[2], [1]], [0.9 * e, 0.0], [3.0, 4.0])
    assert X[0] == 0.0 and X[1] == 0.0 and B[0] == 3.0 and B[1] == 4.0
    # 3) a self slot is skipped, the first zero ends the row (slot 3 unread)
    X, B = _filter([[1, 2, 0, 2], [1, 0, 0, 0]], [1.0, 1.0 + 0.5 * e], [5.0, 6.0])
    assert X[0] == 1.0 + 0.5 * e and B[0] == X[0]
    assert X[1] == X[0] and B[1] == 6.0  # equal to the filtered row 1
    X, B = _filter([[0, 2], [1, 0]], [1.0, 2.0], [5.0, 6.0])
    assert B[0] == 5.0 and X[0] == 1.0 and B[1] == 2.0
    # 4) snaps read the raw value of a higher row and the filtered value of
    #    a lower one: row 2 sees row 1 already moved to row 3's value
    x = [2.0, 2.0 + 0.6 * e, 2.0 + 0.3 * e]
    X, B = _filter([[3, 0], [1, 0], [0, 0]], x, [0.0, 0.0, 0.0])
    assert X[0] == x[2]
    assert X[1] == x[2] and B[1] == x[2]  # |x2 - x3| = 0.3e -> snap
    # 5) beyond the limit nothing moves but B takes X
    X, B = _filter([[2], [1]], [1.0, 1.0 + 1e-12], [0.0, 0.0])
    assert X[0] == 1.0 and X[1] == 1.0 + 1e-12 and B[0] == 1.0 and B[1] == 1.0 + 1e-12


def test_branches_exercised_on_realistic_system():
    from clowns3d import configs
    s = configs.dispersion_2dh(96, 64, seed=43)
    o = call_oracle_will(s, iters=1000)
    assert o["rc"] == 0 and o["IER"] == 0
    xr, x, b = o["X_raw"], o["X"], o["B"]
    z = np.where(np.abs(xr) <= 1e-14, 0.0, xr)
    zero_snapped = (np.abs(xr) <= 1e-14) & (xr != 0)
    moved = x != z  # snapped to a neighbour
    unassigned = b != x
    # chains: a moved row whose value came from a lower row that also moved
    LL = s["LL"]
    chain = 0
    for i in np.nonzero(moved)[0]:
        for k in range(s["NL"]):
            j = LL[i, k] - 1
            if 0 <= j < i and moved[j] and x[j] == x[i]:
                chain += 1
                break
    print(f"N {s['N']} zero-snapped {zero_snapped.sum()} moved {moved.sum()} unassigned {unassigned.sum()} "
          f"chain links {chain}")
    assert zero_snapped.sum() > 0 and moved.sum() > 0 and unassigned.sum() > 0 and chain > 0
