"""Writes the golden fixtures of tests/golden/ from the reference's own CSR5
code (oracle/_ref/libdimref.so, compiled by oracle/Makefile from the reference
tree's unmodified headers).

    make -C oracle && python tests/golden/make_golden.py

Each fixture holds the CSR input, x, the reference's y = A*x (called the way
it is correct: zero-filled y, fresh calibrator) at 1 and 8 OpenMP threads, and
the reference's csr5_info arrays (tile_ptr, packed tile_desc words,
tile_desc_offset_ptr, tile_desc_offset) plus its transposed vals/col_idx.
Only data is stored; no reference source is copied.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from diploma_research_amd import matgen  # noqa: E402
from diploma_research_amd.csr5 import Csr  # noqa: E402
from diploma_research_amd.mmio import read_coo  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cases():
    # the reference's own test matrix (tests/test_files/mm.mtx)
    coo = read_coo(os.path.join(HERE, "mm.mtx"))
    csr = Csr.from_coo(coo)
    yield "mm", coo.dimensions.rows, coo.dimensions.cols, csr.row_start_offsets, csr.col_indices, csr.values, 1
    # empty rows (dirty tiles), long rows, nnz not a multiple of the tile
    rp, ci, v = matgen.random_csr(1000, 1000, seed=7, max_len=9, empty_frac=0.15, long_rows=[(3, 300), (500, 700)])
    yield "rand_dirty", 1000, 1000, rp, ci, v, 2
    # whole tiles inside one row (fast track) and a row spanning many tiles
    rp, ci, v = matgen.random_csr(300, 5000, seed=11, max_len=4, long_rows=[(10, 4000), (11, 64), (200, 1000)])
    yield "rand_longrows", 300, 5000, rp, ci, v, 3
    # rectangular, nnz a multiple of sigma*omega = 64
    rp, ci, v = matgen.random_csr(640, 97, seed=5, max_len=6)
    k = (len(v) // 64) * 64
    rp = np.minimum(rp, k).astype(np.uint32)
    yield "rect_exact", 640, 97, rp, ci[:k], v[:k], 4
    # the 2-D 5-point Laplacian on 256 x 256 (stored as digests + y only)
    rp, ci, v = matgen.poisson2d(256)
    yield "poisson2d_256", 65536, 65536, rp, ci, v, 5


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libdimref.so not built (make -C oracle with the reference tree present)")
    for name, rows, cols, rp, ci, v, seed in cases():
        rng = np.random.default_rng(1000 + seed)
        x = rng.uniform(-1.0, 1.0, cols)
        R = O.RefCsr5(rows, cols, rp, ci, v)
        info = R.info
        y1 = R.spmv(x, nthreads=1)
        y8 = R.spmv(x, nthreads=8)
        big = len(v) > 100000
        out = dict(rows=rows, cols=cols, x=x, y_ref_t1=y1, y_ref_t8=y8, tile_count=info["tile_count"],
                   tail_partition_start=info["tail_partition_start"])
        if big:
            out.update(x_seed=1000 + seed, tile_ptr_sha=_digest(info["tile_ptr"]),
                       tile_desc_sha=_digest(info["tile_desc"]), vals_sha=_digest(info["vals"]),
                       col_idx_sha=_digest(info["col_idx"]))
            out.pop("x")
        else:
            out.update(row_ptr=rp, col_idx=ci, vals=v, tile_ptr=info["tile_ptr"], tile_desc=info["tile_desc"],
                       tile_desc_offset_ptr=info["tile_desc_offset_ptr"],
                       tile_desc_offset=info["tile_desc_offset"], t_vals=info["vals"], t_col_idx=info["col_idx"])
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, rows, cols, len(v), "tiles", info["tile_count"])


if __name__ == "__main__":
    main()
