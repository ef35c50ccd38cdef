"""Time-stepped pressure solves (configuration 5) and the checkpointed adjoint.

Each step t_n = n*dt_step: the surface is the prescribed solitary wave at t_n
(configs.config_solitary), so F, the cell flags, the row numbering and the
surface positions change every step.  Per step: F on the host, host
classification (clowns_classify), upload of what changed (F, flags, numbering),
fused device assembly + scaling + sort, and the solve -- warm-started from the
previous step's pressure carried over cell by cell (a deviation from the
reference, which solves for a correction from zero each pass,
src2LOWNS/1sub-momentum.F90:267-271).

Adjoint over the last K steps with solution checkpointing: the forward run
keeps x_t (per row, on the device) and the step's row-to-cell map; the
reverse sweep re-assembles A_t from the geometry at t (recompute, not store),
solves A_t^T lambda_t = dL/dx_t and evaluates the coefficient gradients with
the checkpointed x_t.  L = sum_t sum_probes w x_t^2 over probe cells.
"""
import time

import numpy as np

from . import configs
from .assembly import DeviceGrid, classify
from .pcgpme import Solver


class SolitaryRun:
    def __init__(self, nx, ny, nz, dt_step=0.05, device=0, seed=2026, on_device=False):
        import torch
        self.torch = torch
        self.on_device = on_device  # F and classification on the GPU
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dt_step = dt_step
        self.seed = seed
        self.dev = torch.device(f"cuda:{device}")
        self.sol = Solver(device)
        self.grid = None
        self.x_cell = None  # previous solution on the box (warm start)

    def geometry(self, n):
        """The grid at step n: static fields built once, F recomputed."""
        t = n * self.dt_step
        if getattr(self, "_g", None) is None:
            self._g = configs.config_solitary(self.nx, self.ny, self.nz, t=t, seed=self.seed)
        else:
            configs.solitary_water(self._g, t)
        return self._g

    def setup_step(self, n):
        """F at t_n and classification: on the host (upload what changed), or
        with on_device=True entirely on the GPU."""
        if self.on_device:
            torch = self.torch
            t0 = time.perf_counter()
            if self.grid is None:
                g = self.geometry(n)
                self.grid = DeviceGrid.on_device(g, self.sol)
            else:
                configs.solitary_fill_device(self.grid.g, self.grid.t["F"], n * self.dt_step)
                self.grid.classify_device(self.sol)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            return self.grid.g, {"gen_s": 0.0, "classify_s": t1 - t0, "upload_s": 0.0,
                                 "host_fallback": self.grid.last_fallback}
        t0 = time.perf_counter()
        g = self.geometry(n)
        t1 = time.perf_counter()
        c = classify(g)
        t2 = time.perf_counter()
        if self.grid is None:
            self.grid = DeviceGrid(g, classified=c)
        else:
            self.grid.update(g, c)
        t3 = time.perf_counter()
        return g, {"gen_s": t1 - t0, "classify_s": t2 - t1, "upload_s": t3 - t2}

    def step(self, n, mixed=True, warm=True, eps=1e-10, keep_raw=False):
        torch = self.torch
        g, tm = self.setup_step(n)
        dg = self.grid
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b = dg.assemble_into(self.sol, keep_raw=keep_raw)
        r2c = dg.row2cell.long()
        use_warm = warm and self.x_cell is not None
        x = self.x_cell[r2c].clone() if use_warm else torch.zeros(dg.N, dtype=torch.float64, device=self.dev)
        x, info = self.sol.solve(b, x, eps=eps, mixed=mixed, warm=use_warm)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if self.x_cell is None:
            self.x_cell = torch.zeros(g.ncell, dtype=torch.float64, device=self.dev)
        self.x_cell.zero_()
        self.x_cell[r2c] = x
        tm.update(device_s=t1 - t0, N=dg.N, iters=info["iters"], refinements=info["refinements"],
                  err=info["err"], ier=info["ier"], warm=use_warm, solve_ms=info["ms"])
        return x, b, tm

    def run(self, steps, mixed=True, warm=True, eps=1e-10, adjoint_steps=0, probes=1000, log=None):
        """Forward `steps` steps; then the adjoint over the last adjoint_steps.
        Returns (per-step records, adjoint records)."""
        torch = self.torch
        recs, ckpt = [], []
        rng = np.random.default_rng(self.seed + 7)
        probe_cells = None
        for n in range(steps):
            keep = n >= steps - adjoint_steps
            x, b, tm = self.step(n, mixed=mixed, warm=warm, eps=eps, keep_raw=keep)
            tm["step"] = n
            recs.append(tm)
            if log:
                log(tm)
            if keep:
                if probe_cells is None:
                    # probes among the rows of the first checkpointed step
                    r2c = self.grid.row2cell
                    pick = rng.choice(self.grid.N, size=min(probes, self.grid.N), replace=False)
                    probe_cells = r2c[torch.as_tensor(pick, device=self.dev)].long()
                    weights = torch.tensor(rng.uniform(0.5, 1.5, len(pick)), device=self.dev)
                ckpt.append({"n": n, "x": x.detach().clone(), "row2cell": self.grid.row2cell.clone()})
        adj = []
        for c in reversed(ckpt):
            n = c["n"]
            g, tm = self.setup_step(n)
            dg = self.grid
            assert torch.equal(dg.row2cell, c["row2cell"]), "re-assembled numbering differs"
            dg.assemble_into(self.sol, keep_raw=True)
            # dL/dx_t on rows: probes -> rows through the cell map
            cell_to_row = torch.full((g.ncell,), -1, dtype=torch.long, device=self.dev)
            cell_to_row[c["row2cell"].long()] = torch.arange(dg.N, device=self.dev)
            rows = cell_to_row[probe_cells]
            ok = rows >= 0
            gx = torch.zeros(dg.N, dtype=torch.float64, device=self.dev)
            gx[rows[ok]] = 2.0 * weights[ok] * c["x"][rows[ok]]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            lam, info = self.sol.solve_transpose(gx, eps=1e-12)
            gAL, gD = self.sol.grad_coefficients(lam, c["x"])
            torch.cuda.synchronize()
            adj.append({"step": n, "N": dg.N, "iters": info["iters"], "ier": info["ier"],
                        "adjoint_s": time.perf_counter() - t0, "grad_b_norm": lam.norm().item(),
                        "grad_AL_norm": gAL.norm().item(), "grad_D_norm": gD.norm().item(),
                        "loss_t": (weights[ok] * c["x"][rows[ok]] ** 2).sum().item()})
        return recs, adj


def main():
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="configuration 5: time-stepped solves + checkpointed adjoint")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--nz", type=int, default=80)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--adjoint", type=int, default=10)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--cold", action="store_true")
    ap.add_argument("--host-classify", action="store_true",
                    help="F and classification on the host (default: on the GPU)")
    a = ap.parse_args()
    run = SolitaryRun(a.nx, a.ny, a.nz, dt_step=a.dt, on_device=not a.host_classify)
    log = lambda r: print(json.dumps(r), flush=True)
    recs, adj = run.run(a.steps, mixed=not a.fp64, warm=not a.cold, adjoint_steps=a.adjoint, log=log)
    for r in adj:
        print(json.dumps(r), flush=True)
    dev_s = [r["device_s"] for r in recs[1:]] or [recs[0]["device_s"]]
    summary = {"grid": [a.nx, a.ny, a.nz], "steps": a.steps, "mixed": not a.fp64, "warm": not a.cold,
               "N_first": recs[0]["N"], "N_last": recs[-1]["N"],
               "device_s_per_step_mean": sum(dev_s) / len(dev_s),
               "iters_mean": sum(r["iters"] for r in recs) / len(recs),
               "host_s_per_step_mean": sum(r["gen_s"] + r["classify_s"] + r["upload_s"] for r in recs) / len(recs),
               "all_converged": all(r["ier"] == 0 for r in recs),
               "adjoint_steps": len(adj), "adjoint_s_mean": (sum(r["adjoint_s"] for r in adj) / len(adj)) if adj else None}
    print("SUMMARY " + json.dumps(summary), flush=True)
    sys.exit(0 if summary["all_converged"] else 1)


if __name__ == "__main__":
    main()
