"""MatrixMarket coordinate reader (the ingest that feeds csr::from_coo).

Follows dim's reader (include/dim/io/matrix_market.h:25-66): coordinate
format only, 1-based indices shifted to 0-based, `pattern` entries read as
1.0, `symmetric` kept as the COO's flag so from_coo mirrors the entries.
"""
from __future__ import annotations

import numpy as np

from .csr5 import Coo, Dimensions


def read_coo(path) -> Coo:
    with open(path, "r") as f:
        banner = f.readline().split()
        if len(banner) < 5 or banner[0] != "%%MatrixMarket" or banner[1].lower() != "matrix":
            raise ValueError("not a MatrixMarket file")
        if banner[2].lower() != "coordinate":
            raise ValueError("only coordinate MatrixMarket files are supported")
        field = banner[3].lower()
        symmetric = banner[4].lower() in ("symmetric", "hermitian", "skew-symmetric")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        rows, cols, nnz = (int(t) for t in line.split()[:3])
        data = np.loadtxt(f, ndmin=2, max_rows=nnz) if nnz else np.zeros((0, 3))
    r = data[:, 0].astype(np.uint32) - 1
    c = data[:, 1].astype(np.uint32) - 1
    v = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    return Coo(Dimensions(rows, cols), v, r, c, symmetric)
