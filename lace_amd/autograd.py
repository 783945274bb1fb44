"""torch.autograd binding of the adjoint over the C ABI -- for testing.

forward : x = A(theta)^{-1} b          (lace_solver_set_values + solve)
backward: A^T lambda = dL/dx           (lace_solver_set_values(theta) of
                                        this node, then lace_solver_solve_trans
                                        with the ILU factors applied transposed)
          dL/dtheta_k = -lambda_i x_j  (lace_solver_grad_values, on A's
                                        pattern), dL/db = lambda

The engine is the C++/HIP library; this wrapper only moves arrays through
host memory.  theta and b are float64 CPU tensors.  One solver may serve
several nodes of a graph: each node keeps its own theta and puts it back
into the solver before its adjoint solve.
"""
from __future__ import annotations

import numpy as np
import torch

from . import DevVec, LaceError, Solver


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, b: torch.Tensor, solver: Solver):
        solver.set_values(theta.detach().cpu().numpy())
        x, st = solver.solve(b.detach().cpu().numpy())
        if st.converged != 1:
            raise LaceError(f"forward solve did not converge (rel {st.rel_true_residual:.2e})")
        xt = torch.from_numpy(x)
        ctx.solver = solver
        ctx.save_for_backward(xt, theta.detach().clone())
        return xt

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        x, theta = ctx.saved_tensors
        S = ctx.solver
        # the solver may hold another node's values by now
        S.set_values(theta.cpu().numpy())
        lam, st = S.solve(gx.detach().cpu().numpy(), trans=1)
        if st.converged != 1:
            raise LaceError(f"adjoint solve did not converge (rel {st.rel_true_residual:.2e})")
        dl = DevVec(S.n, np.float64, lam)
        dx = DevVec(S.n, np.float64, x.numpy())
        nnz = int(S.nnz)
        dtheta = S.grad_values(dl, dx, nnz)
        return torch.from_numpy(dtheta), torch.from_numpy(lam), None


def sparse_solve(theta: torch.Tensor, b: torch.Tensor, solver: Solver) -> torch.Tensor:
    return SparseSolve.apply(theta, b, solver)
