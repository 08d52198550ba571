"""Differentiable projector: torch.autograd.Function over the C-ABI.

forward(x) = A x          (epet_matrix_spmv EPET_OP_N, or epet_matrix_spmm for a
                           (V, F) batch of frame-minor images)
backward(g) = A^T g       (the transposed projection kernel, EPET_OP_T)

so a Poisson-likelihood reconstruction can be written with torch autograd and
checked against MLEM (tests/test_gpu_projector.py).
"""
from __future__ import annotations

import torch

from .api import OP_N, OP_T, Matrix


class Projector(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, A: Matrix) -> torch.Tensor:
        if x.dtype != torch.float32:
            raise TypeError("the projector computes in FP32")
        x = x.contiguous()
        ctx.A = A
        if x.dim() == 1:
            return A.spmv(OP_N, x)
        return A.spmm(OP_N, x)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        A = ctx.A
        g = g.contiguous().float()
        if g.dim() == 1:
            return A.spmv(OP_T, g), None
        return A.spmm(OP_T, g), None


def project(A: Matrix, x: torch.Tensor) -> torch.Tensor:
    """A x with gradient A^T g."""
    return Projector.apply(x, A)


def poisson_nll(A: Matrix, x: torch.Tensor, y: torch.Tensor, eps: float = 1e-8) -> torch.Tensor:
    """sum(A x - y log(A x + eps)), the loss of the differentiable workload."""
    yhat = project(A, x)
    return (yhat - y * torch.log(yhat + eps)).sum()


def projected_gradient_descent(A: Matrix, y: torch.Tensor, x0: torch.Tensor, steps: int,
                               step0: float = None, shrink: float = 0.5):
    """Projected gradient descent x <- max(x - t grad, 0) with Armijo backtracking.
    Returns (x, losses)."""
    x = x0.clone().requires_grad_(True)
    losses = []
    t = step0
    for _ in range(steps):
        loss = poisson_nll(A, x, y)
        (g,) = torch.autograd.grad(loss, x)
        L0 = float(loss)
        losses.append(L0)
        if t is None:
            t = float(x.detach().norm() / (g.norm() + 1e-30))
        with torch.no_grad():
            while True:
                xn = torch.clamp(x - t * g, min=0.0)
                Ln = float(poisson_nll(A, xn, y))
                if Ln <= L0 - 1e-4 / t * float(((xn - x) ** 2).sum()) or t < 1e-30:
                    break
                t *= shrink
            x.copy_(xn)
            t /= shrink   # allow the step to grow again
    with torch.no_grad():
        losses.append(float(poisson_nll(A, x, y)))
    return x.detach(), losses
