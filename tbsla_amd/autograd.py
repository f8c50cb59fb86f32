"""Differentiable PageRank: a torch.autograd.Function over the HIP forward
(tbsla_hip_page_rank_ex) and the hand-written adjoint
(tbsla_hip_page_rank_adjoint).

    p = page_rank(A, values, teleport)          # values: CSR entry order
    loss = (w * p).sum(); loss.backward()       # values.grad, teleport.grad

The forward is the reference PageRank iteration (Matrix.cpp:13-61) run to
its fixed point in FP64, generalised to M = beta*A + (1-beta) v 1^T with the
teleport distribution v (v = 1/n is the reference's uniform teleport).  The
backward solves the implicit adjoint fixed point on the transposed CSR
structure (FP32 products by default) and returns dL/dvalues and dL/dv.
"""
from __future__ import annotations

import torch

from .matrix import MatrixCSR


class PageRankFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, teleport, matrix, beta, epsilon, max_iteration, adjoint):
        n = matrix.get_n_col()
        p = torch.empty(n, dtype=torch.float64, device=values.device)
        matrix.page_rank_device(p, epsilon=epsilon, beta=beta, max_iteration=max_iteration, precision="f64",
                                rule="sumnorm", values=values.detach().contiguous(),
                                teleport=None if teleport is None else teleport.detach().contiguous())
        ctx.matrix, ctx.beta, ctx.adjoint = matrix, beta, dict(adjoint)
        ctx.has_teleport = teleport is not None
        ctx.save_for_backward(values, teleport if teleport is not None else values.new_empty(0), p)
        return p

    @staticmethod
    def backward(ctx, grad_p):
        values, teleport, p = ctx.saved_tensors
        teleport = teleport if ctx.has_teleport else None
        gv = torch.zeros_like(values) if ctx.needs_input_grad[0] else None
        gt = None
        if ctx.needs_input_grad[1]:
            gt = torch.zeros_like(teleport)
        ctx.matrix.page_rank_adjoint(p, grad_p.detach().to(torch.float64).contiguous(), grad_values=gv,
                                     grad_teleport=gt, beta=ctx.beta, values=values.detach().contiguous(),
                                     teleport=teleport, **ctx.adjoint)
        return gv, gt, None, None, None, None, None


def page_rank(matrix: MatrixCSR, values: torch.Tensor, teleport: torch.Tensor | None = None, beta: float = 0.85,
              epsilon: float = 1e-14, max_iteration: int = 1000, adjoint_precision: str = "f32",
              adjoint_epsilon: float = 1e-7, adjoint_max_iteration: int = 1000) -> torch.Tensor:
    """Sum-normalised PageRank of the CSR structure of ``matrix`` with edge
    values ``values`` (float64, one per stored entry) and teleport
    distribution ``teleport`` (float64, n; None = uniform, not
    differentiable).  Differentiable in ``values`` and ``teleport``."""
    adjoint = {"precision": adjoint_precision, "epsilon": adjoint_epsilon,
               "max_iteration": adjoint_max_iteration}
    return PageRankFunction.apply(values, teleport, matrix, beta, epsilon, max_iteration, adjoint)
