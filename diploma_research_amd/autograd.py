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

from . import native as N
from .csr5 import Csr5, conjugate_gradient_batched


def _require_converged(results, what):
    """Autograd callers never see the per-system CgResult, so a solve that
    stopped at max_iters (its x or its gradient would be silently wrong)
    raises instead."""
    bad = [i for i, r in enumerate(results) if not r.converged]
    if bad:
        worst = max(results[i].relres for i in bad)
        raise N.DimError(N.DIM_NOT_CONVERGED,
                         f"{what}: {len(bad)} of {len(results)} systems stopped at max_iters above the tolerance "
                         f"(first {bad[:5]}, largest relative residual {worst:.3e})")


class CsrSolve(torch.autograd.Function):
    """x = solve(A(vals), b) for a (batched) SPD CSR5 matrix handle."""

    @staticmethod
    def forward(ctx, vals, b, A: Csr5, rtol: float, max_iters: int, symmetric: bool = True):
        A.set_values(vals.detach().contiguous())
        x, res = conjugate_gradient_batched(A, b.detach().contiguous(), rtol=rtol, max_iters=max_iters)
        _require_converged(res, "forward solve")
        ctx.A, ctx.rtol, ctx.max_iters, ctx.symmetric = A, rtol, max_iters, symmetric
        ctx.save_for_backward(vals.detach(), x)
        return x

    @staticmethod
    def backward(ctx, gx):
        vals, x = ctx.saved_tensors
        A = ctx.A
        A.set_values(vals.contiguous())  # the handle may have been re-valued since forward
        lam, res = conjugate_gradient_batched(A, gx.contiguous(), rtol=ctx.rtol, max_iters=ctx.max_iters,
                                              transpose=not ctx.symmetric)
        _require_converged(res, "adjoint solve")
        gvals = A.sampled_outer(lam, x, scale=-1.0) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gvals, gb, None, None, None, None


def solve(A: Csr5, vals, b, rtol=1e-12, max_iters=20000, symmetric=True):
    """Differentiable x = A(vals)^-1 b (vals: float64 GPU tensor in CSR order).
    symmetric: the adjoint solve uses A's own product (A^T = A); False: A^T
    products."""
    return CsrSolve.apply(vals, b, A, rtol, max_iters, symmetric)
