"""torch.autograd wrapper of the outlet-flux adjoint, for gradient checks.

OutletFlux.apply(log_c, box) returns the total outlet flux J of the
per-cell coefficient stencil with c = exp(log_c) (an FP64 AMG-CG solve on
the GPU); backward runs the adjoint CG solve with the same hierarchy and the
fused gradient pass through the C ABI.  Torch only carries the tensors; all
work is done by liblineal_amd.
"""

from __future__ import annotations

import numpy as np
import torch

import lineal_amd as la


class OutletFlux(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_c: torch.Tensor, box: "la.BoxInfo", rel_tol: float = 1e-10,
                max_it: int = 1000):
        c = np.exp(log_c.detach().to("cpu", torch.float64).numpy().reshape(-1))
        A = la.StencilMatrix.values(box, c)
        M = la.MultiGridSolver(precision=np.float64).instantiate(A)
        p = la.DenseVector(A.row_num, fill=0.0)
        res = la.ConjugateGradientsSolver(8, M).solve(A, p, A.rhs(), rel_tol, max_it)
        if not res.converged:
            raise la.LinealError(res.status, "outlet-flux forward solve did not converge")
        ctx.state = (A, M, p, rel_tol, max_it)
        ctx.shape, ctx.device = log_c.shape, log_c.device
        return torch.tensor(la.outlet_flux(A, p), dtype=torch.float64, device=log_c.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        A, M, p, rel_tol, max_it = ctx.state
        n = A.row_num
        g = la.outlet_flux_adjoint_rhs(A, la.DenseVector(n))
        lam = la.DenseVector(n, fill=0.0)
        res = la.ConjugateGradientsSolver(8, M).solve(A, lam, g, rel_tol, max_it)
        if not res.converged:
            raise la.LinealError(res.status, "outlet-flux adjoint solve did not converge")
        grad = la.outlet_flux_gradient(A, p, lam, la.DenseVector(n)).numpy().reshape(ctx.shape)
        out = torch.from_numpy(grad).to(ctx.device) * grad_out.to(torch.float64)
        return out, None, None, None


def outlet_flux(log_c: torch.Tensor, box: "la.BoxInfo", rel_tol: float = 1e-10,
                max_it: int = 1000) -> torch.Tensor:
    return OutletFlux.apply(log_c, box, rel_tol, max_it)
