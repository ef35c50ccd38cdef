"""Writes tests/golden/pcgpme_will_*.npz from the flang build of the
reference's PCGPME_WILL (oracle/_ref/libref_pcgpme.so, compiled from
src2LOWNS/pcgpme_2DH.f90 by oracle/Makefile) at one OpenMP thread.

Each fixture holds the inputs (AL, LL, D, B on the 2DH dispersion matrix of
clowns3d.configs.dispersion_2dh) and the reference's B on return for
ITER = 0, 1, 2, 5 and 1000 at EPS = 1e-10, PARM = 5.  Run from the
repository root:
    python tests/golden/make_golden_will.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import conftest  # noqa: E402,F401  (puts the package on sys.path)
from conftest import call_ref_will  # noqa: E402
from clowns3d import configs  # noqa: E402

CASES = {
    "disp40x30_s31": dict(shape=(40, 30), seed=31),
    "disp72x40_s32": dict(shape=(72, 40), seed=32),
}
ITERS = (0, 1, 2, 5, 1000)


def system(case):
    return configs.dispersion_2dh(*case["shape"], seed=case["seed"])


def main():
    for name, case in CASES.items():
        s = system(case)
        out = dict(AL_in=s["AL"], LL_in=s["LL"], D_in=s["D"], B_in=s["B"], N=s["N"], NA=s["NA"], NL=s["NL"])
        for iters in ITERS:
            r = call_ref_will(s, iters=iters, eps=1e-10, nthr=1)
            out[f"it{iters}_B"] = r["B"]
        path = os.path.join(HERE, f"pcgpme_will_{name}.npz")
        np.savez_compressed(path, **out)
        print(path, s["N"])


if __name__ == "__main__":
    main()
