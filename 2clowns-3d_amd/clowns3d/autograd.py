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

OpenRatioSolve (pcg_solve_open_ratios) goes one step further back: x as a
function of a grid's face open ratios a and cell open ratios fb, through the
device assembly of A and b (init_matrix's cal_alu_f, src2LOWNS/1sub-matrix.F90:
50-61,82-87; cal_matrix_element; the RHS).  Its backward is the same
transpose solve followed by the assembly's adjoint
(clowns_grad_open_ratios): dL/da per face and dL/dfb per cell.
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


class OpenRatioSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, fb, grid, solver, eps):
        view = grid.with_open_ratios(a.detach().contiguous(), fb.detach().contiguous())
        b = view.assemble_into(solver, keep_raw=True)
        x, info = solver.solve(b, eps=eps)
        if info["ier"] != 0:
            raise RuntimeError(f"forward solve did not converge: {info}")
        ctx.save_for_backward(a, fb, x)
        ctx.grid, ctx.solver, ctx.eps = grid, solver, eps
        ctx.version = solver.version
        return x

    @staticmethod
    def backward(ctx, gx):
        a, fb, x = ctx.saved_tensors
        solver = ctx.solver
        view = ctx.grid.with_open_ratios(a.detach().contiguous(), fb.detach().contiguous())
        if solver.version != ctx.version:
            view.assemble_into(solver, keep_raw=True)
        lam, info = solver.solve_transpose(gx.contiguous(), eps=ctx.eps)
        if info["ier"] != 0:
            raise RuntimeError(f"adjoint solve did not converge: {info}")
        ga, gfb = view.grad_open_ratios(solver, lam, x)
        return ga, gfb, None, None, None


def pcg_solve_open_ratios(grid, a, fb, solver=None, eps=1e-12):
    """x = A(a, fb)^{-1} b(a, fb) for a DeviceGrid whose face open ratios a
    (the layout of grid.t["a"]) and cell open ratios fb replace its own;
    differentiable in a and fb."""
    if solver is None:
        solver = Solver(a.device.index or 0)
    return OpenRatioSolve.apply(a, fb, grid, solver, eps)


def pcg_solve(AL, D, b, LL, solver=None, eps=1e-12):
    """x = A^{-1} b, differentiable in AL, D and b."""
    if solver is None:
        solver = Solver(b.device.index or 0)
    return PCGSolve.apply(AL, D, b, LL, solver, eps)
