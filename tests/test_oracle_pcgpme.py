"""Pins the C oracle of PCGPME (oracle/oracle_pcgpme.c).

1. Golden fixtures written by the flang build of the reference's own
   src2LOWNS/pcgpme_para.F90 (tests/golden/make_golden_pcgpme.py): every
   output bit for bit at one thread, for ITER = 1, 2, 5 and to convergence.
2. The flang build itself (when oracle/_ref is present), on fresh inputs,
   including NL = 4, a warm start and the EPS/ITER defaults.
3. Mathematics that fixes what the fixtures leave open: AXSUB is a dense
   matrix-vector product and LDUSUB a dense product of (I - L_j) factors of
   the scaled, sorted operator; the converged X solves the dense system;
   the uniform-grid coefficients and slot orders are known in closed form.
"""
import ctypes
import glob
import os

import numpy as np
import pytest

from conftest import (ROOT, call_oracle, call_ref, copy_system, dense_matrix, oracle, ptr)
from _systems import box_system

GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "pcgpme_box*.npz")))


def _from_golden(z):
    return dict(AL=np.asfortranarray(z["AL_in"]), LL=np.asfortranarray(z["LL_in"]), D=z["D_in"].copy(),
                B=z["B_in"].copy(), N=int(z["N"]), NA=int(z["NA"]), NL=int(z["NL"]))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
@pytest.mark.parametrize("iters", [1, 2, 5, 10000])
def test_oracle_matches_reference_golden(path, iters):
    z = np.load(path)
    s = _from_golden(z)
    o = call_oracle(s, iters=iters, nthr=1)
    tag = f"it{iters}"
    for k in ("AL", "LL", "D", "B", "X"):
        assert np.array_equal(o[k], z[f"{tag}_{k}"]), k
    assert np.array_equal(o["EPS"], z[f"{tag}_EPS"])
    assert o["ITER"] == int(z[f"{tag}_ITER"])
    assert o["IER"] == int(z[f"{tag}_IER"])


def test_golden_present():
    assert len(GOLDEN) >= 3


@pytest.mark.parametrize("shape,seed", [((17, 9, 7), 31), ((40, 24, 10), 32)])
def test_oracle_matches_flang_reference(shape, seed):
    s = box_system(*shape, seed=seed)
    for iters in (3, 10000):
        o = call_oracle(s, iters=iters)
        r = call_ref(s, iters=iters)
        for k in ("AL", "LL", "D", "B", "X", "EPS"):
            assert np.array_equal(o[k], r[k]), k
        assert (o["ITER"], o["IER"]) == (r["ITER"], r["IER"])
    # warm start and the EPS <= TINY / ITER <= 0 defaults (pcgpme_para.F90:51-52)
    x0 = np.random.default_rng(1).uniform(-1, 1, s["N"])
    o = call_oracle(s, x0=x0, eps=0.0, iters=0)
    r = call_ref(s, x0=x0, eps=0.0, iters=0)
    assert np.array_equal(o["X"], r["X"]) and (o["ITER"], o["IER"]) == (r["ITER"], r["IER"])


def test_oracle_matches_flang_reference_nl4():
    s = box_system(30, 1, 12, seed=33)
    s["AL"] = np.asfortranarray(s["AL"][:, [0, 1, 4, 5]])
    s["LL"] = np.asfortranarray(s["LL"][:, [0, 1, 4, 5]])
    s["NL"] = 4
    o = call_oracle(s)
    r = call_ref(s)
    assert np.array_equal(o["X"], r["X"]) and o["ITER"] == r["ITER"]


def test_oracle_threads_deterministic():
    """Per-chunk partials combined in chunk order: repeatable per thread count."""
    s = box_system(40, 24, 10, seed=34)
    a = call_oracle(s, nthr=4)
    b = call_oracle(s, nthr=4)
    assert np.array_equal(a["X"], b["X"])


def _scaled_sorted(s):
    c = copy_system(s)
    o = oracle()
    o.oracle_scale_sort(ptr(c["AL"]), c["NL"], c["NA"], ptr(c["LL"]), ptr(c["D"]), c["N"], ptr(c["B"]),
                        ctypes.c_double(5.0))
    return c


def test_axsub_is_dense_product():
    s = box_system(9, 7, 5, seed=35)
    c = _scaled_sorted(s)
    N = c["N"]
    A = np.diag(c["D"])
    for k in range(6):
        for i in range(N):
            j = c["LL"][i, k] - 1
            if j < N:
                A[i, j] += c["AL"][i, k]
            else:
                assert c["AL"][i, k] == 0.0  # ghost slots carry zero
    x = np.concatenate([np.random.default_rng(2).standard_normal(N), np.zeros(64)])
    y = np.zeros(N + 64)
    assert oracle().oracle_axsub(ptr(c["AL"]), 6, c["NA"], ptr(c["LL"]), ptr(c["D"]), N, ptr(x), ptr(y)) == 0
    np.testing.assert_allclose(y[:N], A @ x[:N], rtol=1e-13, atol=1e-14)
    # and the scaled operator is W A with W from the row-scaling rule
    A0 = dense_matrix(s)
    w = c["D"] / s["D"]
    np.testing.assert_allclose(A, w[:, None] * A0, rtol=1e-13, atol=1e-300)


def test_ldusub_is_product_of_factors():
    s = box_system(8, 6, 5, seed=36)
    c = _scaled_sorted(s)
    N = c["N"]
    I = np.eye(N)
    Ls = []
    for k in range(6):
        Lk = np.zeros((N, N))
        for i in range(N):
            j = c["LL"][i, k] - 1
            if j < N:
                Lk[i, j] = c["AL"][i, k]
        Ls.append(Lk)
    M = (I - Ls[4] - Ls[5]) @ (I - Ls[3]) @ (I - Ls[2]) @ (I - Ls[1]) @ (I - Ls[0])
    p = np.concatenate([np.random.default_rng(3).standard_normal(N), np.zeros(64)])
    q = p.copy()
    w = np.zeros(N + 64)
    oracle().oracle_ldusub(ptr(c["AL"]), 6, c["NA"], ptr(c["LL"]), N, ptr(q), ptr(q), ptr(w))
    np.testing.assert_allclose(q[:N], M @ p[:N], rtol=1e-12, atol=1e-14)
    # a transposed factor order (a plausible slip) would not match
    Mbad = (I - Ls[0]) @ (I - Ls[1]) @ (I - Ls[2]) @ (I - Ls[3]) @ (I - Ls[4] - Ls[5])
    assert np.linalg.norm(Mbad @ p[:N] - q[:N]) > 1e-6 * np.linalg.norm(q[:N])


def test_converged_x_solves_dense_system():
    s = box_system(10, 8, 6, seed=37)
    o = call_oracle(s, eps=1e-13)
    A = dense_matrix(s)
    x = np.linalg.solve(A, s["B"])
    assert np.linalg.norm(o["X"][: s["N"]] - x) <= 1e-10 * np.linalg.norm(x)
    # the solution does not depend on the row scaling (a3): scale rows by 3
    s3 = copy_system(s)
    s3["AL"][: s["N"]] *= 3.0
    s3["D"] *= 3.0
    s3["B"] *= 3.0
    o3 = call_oracle(s3, eps=1e-13)
    assert np.linalg.norm(o3["X"] - o["X"]) <= 1e-10 * np.linalg.norm(x)


def test_uniform_grid_known_answer():
    """dx=dy=1, dz=0.5, a=fb=1: mp=0.25, ALU_x=-0.25, ALU_z=-1, D=3 before
    scaling; W=1/3 after, AL_x=-1/12, AL_z=-1/3, D=1; interior slot order
    [z-, z+, y-, y+, x-, x+]; a row with x- and z- closed sorts to
    [z+, y-, y+, x+, z-, x-] (the strict-< exchange sort is not stable)."""
    s = box_system(6, 6, 6, seed=0, lo=1.0)  # all a = fb = 1
    N, nx, ny = s["N"], 6, 6
    row = lambda i, j, k: i + nx * (j + ny * k)
    r = row(3, 3, 2)
    assert s["D"][r] == 3.0
    assert list(s["AL"][r]) == [-0.25, -0.25, -0.25, -0.25, -1.0, -1.0]
    c = _scaled_sorted(s)
    assert c["D"][r] == 1.0
    assert np.allclose(c["AL"][r], [-1 / 3, -1 / 3, -1 / 12, -1 / 12, -1 / 12, -1 / 12], rtol=0, atol=1e-16)
    nb = c["LL"][r] - 1
    assert list(nb) == [r - nx * ny, r + nx * ny, r - nx, r + nx, r - 1, r + 1]
    # x- and z- closed: interior in y, wall at i=0, bottom k=0
    r2 = row(0, 3, 0)
    nb2 = c["LL"][r2] - 1
    expect = [r2 + nx * ny, r2 - nx, r2 + nx, r2 + 1]
    assert list(nb2[:4]) == expect
    assert all(v >= N for v in nb2[4:])  # the two ghost slots, in z-, x- order
    assert nb2[4] == N + (r2 + 1) % 64 and nb2[5] == N + (r2 + 1) % 64
