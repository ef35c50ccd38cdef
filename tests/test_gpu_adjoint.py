"""GPU: the hand-written adjoint.

1. Small operator: gradients from PCGSolve.backward (transpose solve +
   gradient kernel) against the FP64 dense reference gradient
   (numpy: lambda = A^-T dL/dx, dL/dA = -lambda x^T) to 1e-8 relative.
2. Configuration 4: the configuration-2 operator at 256x256x64,
   L = sum_i w_i x_i^2 over 1000 seeded probe rows, w ~ U(0.5, 1.5);
   directional derivatives along 20 seeded random directions in (AL, D, b)
   (off-diagonals perturbed only where LL != 0) against FP64 central
   differences with eps = 1e-6, to 1e-6 relative.  Each perturbed solve is
   done as a correction x0 + A'^{-1}(b' - A' x0) with the residual formed in
   FP64, so solver tolerance does not swamp the O(eps) difference.
3. The same loss differentiated with respect to the face open ratios a and
   the cell open ratios fb through the assembly (pcg_solve_open_ratios:
   transpose solve + clowns_grad_open_ratios), against central differences
   of the reassembled system along 20 seeded directions in (a, fb), at
   configuration 4 and on a small bathymetry grid with partial cells.
"""
import numpy as np
import pytest
import torch

from clowns3d import DeviceGrid, Solver, configs, pcg_solve, pcg_solve_open_ratios
from conftest import dense_matrix
from _systems import box_system

pytestmark = pytest.mark.gpu
dev = torch.device("cuda:0")


def test_adjoint_matches_dense_gradient():
    s = box_system(10, 8, 6, seed=41)
    N = s["N"]
    AL = torch.tensor(np.ascontiguousarray(s["AL"].T), device=dev, requires_grad=True)
    LL = torch.tensor(np.ascontiguousarray(s["LL"].T), device=dev)
    D = torch.tensor(s["D"], device=dev, requires_grad=True)
    b = torch.tensor(s["B"], device=dev, requires_grad=True)
    w = torch.linspace(0.5, 1.5, N, dtype=torch.float64, device=dev)
    x = pcg_solve(AL, D, b, LL, eps=1e-13)
    L = (w * x * x).sum()
    L.backward()
    A = dense_matrix(s)
    xd = np.linalg.solve(A, s["B"])
    lam = np.linalg.solve(A.T, 2 * w.cpu().numpy() * xd)
    gb = lam
    gD = -lam * xd
    gAL = np.zeros_like(s["AL"].T)
    for k in range(6):
        m = s["LL"][:N, k] > 0
        gAL[k, :N][m] = -lam[m] * xd[s["LL"][:N, k][m] - 1]
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)
    assert rel(x.detach().cpu().numpy(), xd) < 1e-10
    assert rel(b.grad.cpu().numpy(), gb) < 1e-8
    assert rel(D.grad.cpu().numpy(), gD) < 1e-8
    assert rel(AL.grad.cpu().numpy(), gAL) < 1e-8


def _apply(AL, LL, D, x):
    """y = A x on the device, reference layout (AL, LL: (6, lda))."""
    N = D.shape[0]
    y = D * x
    for k in range(AL.shape[0]):
        idx = LL[k, :N].long()
        m = idx > 0
        y = y + torch.where(m, AL[k, :N] * x[(idx - 1).clamp(min=0)], torch.zeros_like(y))
    return y


def test_config4_adjoint_vs_central_differences():
    g = configs.config2()
    dg = DeviceGrid(g)
    AL, LL, D, b = dg.assemble()
    N = dg.N
    rng = np.random.default_rng(4)
    probes = torch.tensor(rng.choice(N, 1000, replace=False), device=dev)
    w = torch.tensor(rng.uniform(0.5, 1.5, 1000), device=dev)
    sol = Solver(0)

    ALr = AL.clone().requires_grad_(True)
    Dr = D.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    x0 = pcg_solve(ALr, Dr, br, LL, solver=sol, eps=1e-12)
    L0 = (w * x0[probes] ** 2).sum()
    L0.backward()
    gAL, gD, gb = ALr.grad, Dr.grad, br.grad
    x0 = x0.detach()

    def loss_at(ALp, Dp, bp):
        r = bp - _apply(ALp, LL, Dp, x0)
        sol.set_matrix(ALp.contiguous(), LL, Dp.contiguous())
        d, info = sol.solve(r.contiguous(), eps=1e-12)
        assert info["ier"] == 0, info
        x = x0 + d
        return (w * x[probes] ** 2).sum().item()

    mask = (LL[:, :N] > 0).to(torch.float64)
    eps = 1e-6
    errs = []
    gen = torch.Generator(device="cpu").manual_seed(20)
    for t in range(20):
        vAL = torch.zeros_like(AL)
        vAL[:, :N] = torch.randn(AL[:, :N].shape, generator=gen, dtype=torch.float64).to(dev) * mask * AL[:, :N].abs()
        vD = torch.randn(N, generator=gen, dtype=torch.float64).to(dev) * D.abs()
        vb = torch.randn(N, generator=gen, dtype=torch.float64).to(dev) * b.abs().mean()
        pred = ((gAL * vAL).sum() + (gD * vD).sum() + (gb * vb).sum()).item()
        Lp = loss_at(AL + eps * vAL, D + eps * vD, b + eps * vb)
        Lm = loss_at(AL - eps * vAL, D - eps * vD, b - eps * vb)
        fd = (Lp - Lm) / (2 * eps)
        errs.append(abs(fd - pred) / max(abs(fd), 1e-300))
    print("config 4: N", N, "max rel err", max(errs), "median", float(np.median(errs)))
    assert max(errs) <= 1e-6


def test_solver_orders_with_torch_default_stream():
    """Regression: the solver must run on torch's stream (the null stream by
    default); a zero-fill issued by torch just before a library kernel must
    not land after it."""
    s = box_system(40, 30, 12, seed=43)
    N = s["N"]
    sol = Solver(0)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    sol.set_matrix(t(s["AL"].T), t(s["LL"].T), t(s["D"]))
    x, _ = sol.solve(t(s["B"]), eps=1e-12)
    lam, _ = sol.solve_transpose(t(np.ones(N)), eps=1e-12)
    for _ in range(20):
        gAL, gD = sol.grad_coefficients(lam, x)  # allocates with torch.zeros, then the kernel writes
        assert torch.equal(gD, -lam * x)


def _open_ratio_fd(g, nprobe, ndir, seed):
    dg = DeviceGrid(g)
    N = dg.N
    rng = np.random.default_rng(seed)
    probes = torch.tensor(rng.choice(N, min(nprobe, N), replace=False), device=dev)
    w = torch.tensor(rng.uniform(0.5, 1.5, probes.numel()), device=dev)
    sol = Solver(0)
    a0, fb0 = dg.t["a"].clone(), dg.t["fb"].clone()
    ar, fbr = a0.clone().requires_grad_(True), fb0.clone().requires_grad_(True)
    x0 = pcg_solve_open_ratios(dg, ar, fbr, solver=sol, eps=1e-12)
    (w * x0[probes] ** 2).sum().backward()
    ga, gfb = ar.grad, fbr.grad
    x0 = x0.detach()
    assert ga.abs().sum() > 0 and gfb.abs().sum() > 0

    def loss_at(a, fb):
        AL, LL, D, b = dg.with_open_ratios(a.contiguous(), fb.contiguous()).assemble()
        r = b - _apply(AL, LL, D, x0)
        sol.set_matrix(AL.contiguous(), LL, D.contiguous())
        d, info = sol.solve(r.contiguous(), eps=1e-12)
        assert info["ier"] == 0, info
        x = x0 + d
        return (w * x[probes] ** 2).sum().item()

    eps = 1e-6
    errs = []
    gen = torch.Generator(device="cpu").manual_seed(seed)
    for _ in range(ndir):
        # closed faces (a = 0) and solid cells stay as they are
        va = torch.randn(a0.shape, generator=gen, dtype=torch.float64).to(dev) * a0
        vfb = torch.randn(fb0.shape, generator=gen, dtype=torch.float64).to(dev) * fb0
        pred = ((ga * va).sum() + (gfb * vfb).sum()).item()
        fd = (loss_at(a0 + eps * va, fb0 + eps * vfb) - loss_at(a0 - eps * va, fb0 - eps * vfb)) / (2 * eps)
        errs.append(abs(fd - pred) / max(abs(fd), 1e-300))
    sol.close()
    return N, errs


def test_open_ratio_gradient_partial_cells():
    """Bathymetry grid with partial cells (fb < 1 on 1832 cells, a < 1 on
    3580 faces)."""
    g = configs.config_bathymetry(48, 40, 16, nbumps=6, seed=7, width=(3, 10))
    N, errs = _open_ratio_fd(g, 500, 20, 31)
    print("open ratios, partial cells: N", N, "max rel err", max(errs))
    assert max(errs) <= 1e-6


def test_config4_open_ratio_gradient_vs_central_differences():
    N, errs = _open_ratio_fd(configs.config2(), 1000, 20, 32)
    print("open ratios, config 4: N", N, "max rel err", max(errs), "median", float(np.median(errs)))
    assert max(errs) <= 1e-6
