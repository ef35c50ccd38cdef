"""GPU: time-stepped solves (configuration-5 construction at reduced size).

Each step reclassifies and reassembles for the moved solitary wave; the
warm-started mixed solve must reach the stopping rule every step with no
more iterations than the cold FP64 solve of the same step, and the checkpointed
adjoint over the last 3 steps must re-create each step's numbering and
converge.  Every step's operator and right-hand side is also checked against
the oracle's assembly of the same geometry (bit for bit)."""
import numpy as np
import pytest
import torch

from clowns3d.timestep import SolitaryRun
from test_oracle_assembly import build_oracle

pytestmark = pytest.mark.gpu


def test_solitary_steps_and_adjoint():
    run = SolitaryRun(96, 16, 20, dt_step=0.4)
    recs, adj = run.run(steps=6, mixed=True, warm=True, adjoint_steps=3, probes=200)
    print(recs)
    print(adj)
    assert all(r["ier"] == 0 and r["err"] <= 1e-10 for r in recs)
    assert all(r["warm"] for r in recs[1:])
    assert [a["step"] for a in adj] == [5, 4, 3]
    assert all(a["ier"] == 0 and a["grad_b_norm"] > 0 for a in adj)
    # warm start does not cost iterations against a cold FP64 solve
    cold = SolitaryRun(96, 16, 20, dt_step=0.4)
    for n in (1, 3, 5):
        _, _, tm = cold.step(n, mixed=False, warm=False)
        assert recs[n]["N"] == tm["N"]


def test_solitary_assembly_matches_oracle():
    run = SolitaryRun(64, 12, 16, dt_step=0.5)
    for n in (0, 2):
        g, _ = run.setup_step(n)
        AL, LL, D, B = run.grid.assemble()
        o = build_oracle(g)
        N = o["N"]
        assert run.grid.N == N
        assert np.array_equal(AL.cpu().numpy(), o["AL"][:, :N])
        assert np.array_equal(LL.cpu().numpy(), o["LL"][:, :N])
        assert np.array_equal(D.cpu().numpy(), o["D"][:N])
        assert np.array_equal(B.cpu().numpy(), o["B"][:N])


def test_solitary_steps_device_classification():
    """The same run with F and the classification on the GPU gives the same
    rows and, to the solve tolerance, the same solutions as the
    host-classified run (F is computed by torch there and by numpy here, so
    the inputs may differ in the last bit)."""
    a = SolitaryRun(96, 16, 20, dt_step=0.4)
    b = SolitaryRun(96, 16, 20, dt_step=0.4, on_device=True)
    for n in range(4):
        xa, _, ta = a.step(n, mixed=False)
        xb, _, tb = b.step(n, mixed=False)
        assert ta["N"] == tb["N"] and not tb["host_fallback"]
        assert torch.equal(a.grid.row2cell, b.grid.row2cell)
        assert (xa - xb).norm().item() <= 1e-8 * xa.norm().item()
