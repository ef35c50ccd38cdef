"""GPU: the solver times its kernel forms once per operator size and keeps
the decision for operators within 1/16 of that size (a time-stepped run
whose row count moves by a few rows does not time again), times again for a
size outside it, and reports the forms it chose (clowns_solver_tuning)."""
import os

import numpy as np
import pytest

from _systems import box_system

pytestmark = pytest.mark.gpu


def test_tuning_kept_for_nearby_sizes(capfd):
    import torch
    from clowns3d import Solver
    os.environ["CLOWNS_VERBOSE"] = "1"
    try:
        sol = Solver(0)  # the switch is read at creation
    finally:
        del os.environ["CLOWNS_VERBOSE"]
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    timings = []
    tunes = []
    for shape in [(160, 128, 6), (161, 128, 6), (160, 128, 8)]:
        s = box_system(*shape, seed=5)
        sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
        x, info = sol.solve(t(s["B"]), eps=1e-10)
        assert info["ier"] == 0
        err = capfd.readouterr().err
        timings.append(err.count("two iterations"))
        tunes.append(sol.tuning())
    sol.close()
    os.environ["CLOWNS_VERBOSE"] = "0"
    try:
        Solver(0).close()  # back to quiet for the solvers that follow
    finally:
        del os.environ["CLOWNS_VERBOSE"]
    print(timings, tunes)
    assert timings == [1, 0, 1]  # 122880 rows; 123648 (+0.6%) kept; 163840 (+33%) timed again
    assert tunes[0] == tunes[1]
    for tn in tunes:
        assert tn["fuse64"] in (0, 1) and tn["stream64"] in (0, 1, 2, 3)
