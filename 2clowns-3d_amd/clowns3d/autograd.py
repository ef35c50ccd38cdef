"""torch.autograd.Function over the solver's C-ABI (gradient checks only).

Forward: x = A^{-1} b with A given by the reference-layout coefficients
(AL off-diagonals per row, D diagonal; LL the fixed pattern) -- the GPU
PM-Bi-CGSTAB solve.  Backward: the hand-written adjoint -- one transpose
solve lambda = A^{-T} dL/dx (clowns_solver_solve_transpose) and the gradient
kernel (clowns_grad_coefficients):
    dL/db = lambda,  dL/dAL(i,k) = -lambda_i x(LL(i,k)),  dL/dD_i = -lambda_i x_i.

Each node owns its solver's resident operator between forward and backward;
if the solver was given another operator in between, backward reloads the
saved one.
"""
import torch

from .pcgpme import Solver


class PCGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, AL, D, b, LL, solver, eps):
        solver.set_matrix(AL.detach().contiguous(), LL, D.detach().contiguous())
        x, info = solver.solve(b.detach().contiguous(), eps=eps)
        if info["ier"] != 0:
            raise RuntimeError(f"forward solve did not converge: {info}")
        ctx.save_for_backward(AL, D, x)
        ctx.LL, ctx.solver, ctx.eps = LL, solver, eps
        ctx.version = solver.version
        return x

    @staticmethod
    def backward(ctx, gx):
        AL, D, x = ctx.saved_tensors
        solver = ctx.solver
        if solver.version != ctx.version:
            solver.set_matrix(AL.detach().contiguous(), ctx.LL, D.detach().contiguous())
        lam, info = solver.solve_transpose(gx.contiguous(), eps=ctx.eps)
        if info["ier"] != 0:
            raise RuntimeError(f"adjoint solve did not converge: {info}")
        gAL, gD = solver.grad_coefficients(lam, x)
        return gAL, gD, lam, None, None, None


def pcg_solve(AL, D, b, LL, solver=None, eps=1e-12):
    """x = A^{-1} b, differentiable in AL, D and b."""
    if solver is None:
        solver = Solver(b.device.index or 0)
    return PCGSolve.apply(AL, D, b, LL, solver, eps)
