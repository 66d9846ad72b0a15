"""Differentiable CG solve behind torch.autograd, over the same C ABI.

x = A(theta)^-1 b, with theta the stored values of A in CSR order. For a loss
L(x) the hand-written adjoint is

    A^T lambda = dL/dx                (adjoint CG: the batched solver)
    dL/db      = lambda
    dL/dA_ij   = -lambda_i x_j        (sampled outer product on the sparsity pattern)

CG needs A symmetric positive definite, so A^T = A at the point where the
gradient is taken, and the adjoint solve runs the same product as the forward
solve (deterministic, no atomics). symmetric=False solves with A^T products
(the transposed kernel, with atomic adds) instead.

Gradients are with respect to every stored entry independently; for a
symmetric parametrisation (a_ij and a_ji tied) add the two entries' gradients.
"""
from __future__ import annotations

import torch

from .csr5 import Csr5, conjugate_gradient_batched


class CsrSolve(torch.autograd.Function):
    """x = solve(A(vals), b) for a (batched) SPD CSR5 matrix handle."""

    @staticmethod
    def forward(ctx, vals, b, A: Csr5, rtol: float, max_iters: int, symmetric: bool = True):
        A.set_values(vals.detach().contiguous())
        x, res = conjugate_gradient_batched(A, b.detach().contiguous(), rtol=rtol, max_iters=max_iters)
        ctx.A, ctx.rtol, ctx.max_iters, ctx.symmetric = A, rtol, max_iters, symmetric
        ctx.results = res
        ctx.save_for_backward(vals.detach(), x)
        return x

    @staticmethod
    def backward(ctx, gx):
        vals, x = ctx.saved_tensors
        A = ctx.A
        A.set_values(vals.contiguous())  # the handle may have been re-valued since forward
        lam, _ = conjugate_gradient_batched(A, gx.contiguous(), rtol=ctx.rtol, max_iters=ctx.max_iters,
                                            transpose=not ctx.symmetric)
        gvals = A.sampled_outer(lam, x, scale=-1.0) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gvals, gb, None, None, None, None


def solve(A: Csr5, vals, b, rtol=1e-12, max_iters=20000, symmetric=True):
    """Differentiable x = A(vals)^-1 b (vals: float64 GPU tensor in CSR order).
    symmetric: the adjoint solve uses A's own product (A^T = A); False: A^T
    products."""
    return CsrSolve.apply(vals, b, A, rtol, max_iters, symmetric)
