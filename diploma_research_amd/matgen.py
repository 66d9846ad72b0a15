"""Seeded synthetic matrices of the CSR5 benchmark workloads (CSR, uint32 indices).

All generators return (row_ptr, col_idx, vals) as numpy arrays with sorted
columns in every row, matching dim's csr container (include/dim/mat/
storage_formats/csr.h:19-78).
"""
from __future__ import annotations

import numpy as np


def _stencil_csr(shape, diag, offdiag_fn):
    """Assembles a Dirichlet stencil operator on a structured grid.

    diag: (n,) diagonal values; offdiag_fn(axis, lo_idx) -> coefficients of the
    couplings between cell lo_idx and its +1 neighbour along `axis` (negated in
    the matrix). Missing neighbours at the boundary are dropped.
    """
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.int64).reshape(shape)
    rows = [np.arange(n, dtype=np.int64)]
    cols = [np.arange(n, dtype=np.int64)]
    vals = [np.asarray(diag, dtype=np.float64)]
    for ax in range(len(shape)):
        lo = np.take(idx, np.arange(shape[ax] - 1), axis=ax).ravel()
        hi = np.take(idx, np.arange(1, shape[ax]), axis=ax).ravel()
        w = offdiag_fn(ax, lo, hi)
        rows += [lo, hi]
        cols += [hi, lo]
        vals += [-w, -w]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    row_ptr = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(np.bincount(r, minlength=n), out=row_ptr[1:])
    return row_ptr, c.astype(np.uint32), v


def poisson2d(side):
    """5-point Laplacian on side x side (diagonal 4, neighbours -1, Dirichlet)."""
    n = side * side
    return _stencil_csr((side, side), np.full(n, 4.0), lambda ax, lo, hi: np.ones(len(lo)))


def poisson3d(side):
    """7-point Laplacian on side^3 (diagonal 6, neighbours -1, Dirichlet)."""
    n = side ** 3
    return _stencil_csr((side, side, side), np.full(n, 6.0), lambda ax, lo, hi: np.ones(len(lo)))


def varcoef(shape, sigma_log, seed):
    """Variable-coefficient diffusion operator on a 2-D or 3-D grid.

    Cell diffusivity kappa ~ lognormal(0, sigma_log); interior faces use the
    harmonic mean 2 k1 k2 / (k1 + k2); boundary faces use the cell's kappa
    (Dirichlet). The diagonal is the sum of the cell's face coefficients, so A
    is SPD and strictly diagonally dominant on boundary rows.
    """
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    kappa = rng.lognormal(0.0, sigma_log, size=n)
    kap = kappa.reshape(shape)
    diag = np.zeros(n)
    face = {}
    for ax in range(len(shape)):
        k1 = np.take(kap, np.arange(shape[ax] - 1), axis=ax)
        k2 = np.take(kap, np.arange(1, shape[ax]), axis=ax)
        h = 2.0 * k1 * k2 / (k1 + k2)
        face[ax] = h.ravel()
        d = np.zeros(shape)
        sl_lo = [slice(None)] * len(shape)
        sl_hi = [slice(None)] * len(shape)
        sl_lo[ax] = slice(0, shape[ax] - 1)
        sl_hi[ax] = slice(1, shape[ax])
        d[tuple(sl_lo)] += h
        d[tuple(sl_hi)] += h
        # two boundary faces per axis
        b_lo = [slice(None)] * len(shape)
        b_hi = [slice(None)] * len(shape)
        b_lo[ax] = slice(0, 1)
        b_hi[ax] = slice(shape[ax] - 1, shape[ax])
        d[tuple(b_lo)] += kap[tuple(b_lo)]
        d[tuple(b_hi)] += kap[tuple(b_hi)]
        diag += d.ravel()
    return _stencil_csr(tuple(shape), diag, lambda ax, lo, hi: face[ax])


def powerlaw(n, seed, alpha=2.0, lo=1, hi=10000, xm=8.0, dtype=np.float64):
    """Irregular-row matrix: row lengths from a Pareto law of shape alpha
    (P(len > x) = (xm / x)^alpha) with scale xm, floored and clipped to
    [lo, hi]: mean about 15.5 for alpha=2, xm=8. Columns uniform without
    duplicates in a row (sorted); values N(0, 1)."""
    rng = np.random.default_rng(seed)
    u = 1.0 - rng.random(n)  # (0, 1]
    lens = np.floor(xm / u ** (1.0 / alpha)).astype(np.int64)
    lens = np.clip(lens, lo, min(hi, n))
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=row_ptr[1:])
    nnz = int(row_ptr[-1])
    # (row, col) keys sorted; duplicates inside a row are redrawn until none remain
    key = np.repeat(np.arange(n, dtype=np.int64), lens) * n + rng.integers(0, n, size=nnz, dtype=np.int64)
    key.sort()
    for _ in range(64):
        dup = np.nonzero(key[1:] == key[:-1])[0] + 1
        if not len(dup):
            break
        key[dup] = (key[dup] // n) * n + rng.integers(0, n, size=len(dup), dtype=np.int64)
        key.sort()
    else:
        raise RuntimeError("could not draw duplicate-free rows")
    cols = (key % n).astype(np.uint32)
    del key
    vals = rng.standard_normal(nnz).astype(dtype)
    return row_ptr.astype(np.uint32), cols, vals


def random_csr(rows, cols, seed, max_len=9, empty_frac=0.0, long_rows=(), integer=False):
    """Small random CSR with optional empty rows and long rows (test inputs)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, size=rows)
    if empty_frac:
        lens[rng.random(rows) < empty_frac] = 0
    for r, L in long_rows:
        lens[r] = min(L, cols)
    row_ptr = np.zeros(rows + 1, dtype=np.uint32)
    np.cumsum(lens, out=row_ptr[1:])
    col_idx = np.concatenate([np.sort(rng.choice(cols, int(L), replace=False)) for L in lens]
                             or [np.zeros(0, np.int64)]).astype(np.uint32)
    if integer:
        vals = rng.integers(-8, 9, size=len(col_idx)).astype(np.float64)
    else:
        vals = rng.standard_normal(len(col_idx))
    return row_ptr, col_idx, vals


def spmv_bytes(nnz, rows, cols, value_bytes=8, index_bytes=4, k=1):
    """Algorithmic bytes of one (multi-RHS) product: values + column indices
    once, x and y once per right-hand side."""
    return (value_bytes + index_bytes) * nnz + k * value_bytes * (rows + cols)
