"""Writes tests/golden/pcgpme_*.npz from the flang build of the reference's
PCGPME (oracle/_ref/libref_pcgpme.so, compiled from
src2LOWNS/pcgpme_para.F90 by oracle/Makefile) at one OpenMP thread.

Each fixture holds the inputs (AL, LL, D, B, X0, EPS, ITER) and the
reference's outputs (AL, LL, D, B, X, EPS, ITER, IER) for ITER = 1, 2, 5 and
to convergence.  Run from the repository root:
    python tests/golden/make_golden_pcgpme.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import call_ref  # noqa: E402
from _systems import box_system  # noqa: E402

CASES = {
    "box8x6x5_s21": dict(shape=(8, 6, 5), seed=21, nl=6),
    "box11x7x6_s22": dict(shape=(11, 7, 6), seed=22, nl=6),
    "box2dv14x1x9_s23": dict(shape=(14, 1, 9), seed=23, nl=4),
}


def system(case):
    s = box_system(*case["shape"], seed=case["seed"])
    if case["nl"] == 4:
        s["AL"] = np.asfortranarray(s["AL"][:, [0, 1, 4, 5]])
        s["LL"] = np.asfortranarray(s["LL"][:, [0, 1, 4, 5]])
        s["NL"] = 4
    return s


def main():
    for name, case in CASES.items():
        s = system(case)
        out = dict(AL_in=s["AL"], LL_in=s["LL"], D_in=s["D"], B_in=s["B"], N=s["N"], NA=s["NA"], NL=s["NL"])
        for iters in (1, 2, 5, 10000):
            r = call_ref(s, iters=iters, eps=1e-10, nthr=1)
            tag = f"it{iters}"
            for k in ("AL", "LL", "D", "B", "X"):
                out[f"{tag}_{k}"] = r[k]
            out[f"{tag}_EPS"] = r["EPS"]
            out[f"{tag}_ITER"] = r["ITER"]
            out[f"{tag}_IER"] = r["IER"]
        path = os.path.join(HERE, f"pcgpme_{name}.npz")
        np.savez_compressed(path, **out)
        print(path, s["N"], out["it10000_ITER"], out["it10000_IER"])


if __name__ == "__main__":
    main()
