"""GPU: configuration 5 at full size -- a 1024x512x80 box (41.9M cells,
about 26M rows) stepped through 100 time steps of the prescribed solitary
wave with F and the classification on the GPU and warm-started mixed
solves, then the checkpointed adjoint over the last 10 steps.

The oracle's single-thread solve at this size takes far longer than a test,
so the bars are what holds at any size:
- step 0's operator and right-hand side bit for bit against the oracle's
  assembly of the same geometry (host build);
- every step: IER 0, sqrt(RN/BN) <= 1e-10, no host fallback of the device
  classification, a warm start from step 1 on;
- one more step's solution against its own system, independently of the
  solver: ||b - A x|| / ||b|| of the assembled reference-layout system
  (AL, LL, D, B, with A x formed in torch from the 1-based LL) at most
  1e-9 (the stopping rule on the diagonally scaled system, D in [1, 3],
  leaves some 3e-11 on the grids where the oracle can check it);
- the adjoint re-creates each step's numbering (asserted inside the run)
  and every transpose solve converges."""
import time

import numpy as np
import pytest
import torch

from test_oracle_assembly import build_oracle

pytestmark = pytest.mark.gpu

SHAPE = (1024, 512, 80)


def _log(t0):
    def log(r):
        if r["step"] % 10 == 0:
            print(f"  step {r['step']}: N {r['N']}, {r['iters']} inner iterations, err {r['err']:.2e}, "
                  f"{time.time() - t0:.1f} s", flush=True)
    return log


def test_config5_step0_assembly_matches_oracle():
    from clowns3d.timestep import SolitaryRun
    t0 = time.time()
    run = SolitaryRun(*SHAPE, dt_step=0.05, on_device=True)
    g, tm = run.setup_step(0)
    assert not tm["host_fallback"]
    AL, LL, D, B = run.grid.assemble()
    torch.cuda.synchronize()
    print(f"config 5 step 0: N {run.grid.N} assembled on the device, {time.time() - t0:.1f} s", flush=True)
    o = build_oracle(g)
    N = o["N"]
    print(f"  oracle assembly done, {time.time() - t0:.1f} s", flush=True)
    assert run.grid.N == N
    assert np.array_equal(run.grid.row2cell.cpu().numpy(), o["row2cell"][:N])
    assert np.array_equal(LL.cpu().numpy(), o["LL"][:, :N])
    assert np.array_equal(AL.cpu().numpy(), o["AL"][:, :N])
    assert np.array_equal(D.cpu().numpy(), o["D"][:N])
    assert np.array_equal(B.cpu().numpy(), o["B"][:N])
    run.sol.close()


def test_config5_100_steps_and_adjoint():
    from clowns3d.timestep import SolitaryRun
    t0 = time.time()
    run = SolitaryRun(*SHAPE, dt_step=0.05, on_device=True)
    recs, adj = run.run(100, mixed=True, warm=True, adjoint_steps=10, log=_log(t0))
    iters = [r["iters"] for r in recs]
    print(f"config 5: N {recs[0]['N']}..{recs[-1]['N']}, inner iterations mean {np.mean(iters):.1f} "
          f"(range {min(iters)}-{max(iters)}), step ms mean {1e3 * np.mean([r['device_s'] for r in recs[1:]]):.1f}",
          flush=True)
    assert recs[0]["N"] > 25_000_000
    assert all(r["ier"] == 0 and r["err"] <= 1e-10 for r in recs)
    assert not any(r["host_fallback"] for r in recs)
    assert not recs[0]["warm"] and all(r["warm"] for r in recs[1:])
    assert [a["step"] for a in adj] == list(range(99, 89, -1))
    assert all(a["ier"] == 0 and a["grad_b_norm"] > 0 and a["grad_AL_norm"] > 0 for a in adj)
    print(f"  adjoint: {len(adj)} steps, iterations {[a['iters'] for a in adj]}, "
          f"{time.time() - t0:.1f} s", flush=True)

    # one more step, checked against its own assembled system
    x, b, tm = run.step(100, mixed=True, warm=True)
    assert tm["ier"] == 0 and tm["err"] <= 1e-10
    AL, LL, D, B = run.grid.assemble()
    assert torch.equal(b, B)
    Ax = D * x
    for k in range(6):
        col = LL[k].long()
        m = col > 0
        Ax[m] += AL[k][m] * x[col[m] - 1]
    res = ((B - Ax).norm() / B.norm()).item()
    print(f"  step 100: N {tm['N']}, {tm['iters']} inner iterations, ||b - A x|| / ||b|| = {res:.3e}", flush=True)
    assert res <= 1e-9
    run.sol.close()
