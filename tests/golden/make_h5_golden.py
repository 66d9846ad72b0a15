"""Writes the HDF5 fixtures of tests/golden/ with the reference's own HDF5
writers (oracle/_ref/libdimref_h5.so, compiled by oracle/Makefile from the
reference's libs/dim/io sources):

    make -C oracle && python tests/golden/make_h5_golden.py

  mm.csr5.h5     group A: the reference's CSR5 group of tests/test_files/mm.mtx
                 (h5::store(csr5::from_csr(...)), libs/dim/io/h5.cpp:45-55)
  mm.csr.h5      group A: its MATLAB CSR group (write_matlab_compatible)
  dirty.csr5.h5  group B: a random 900 x 700 matrix with empty rows (dirty
                 tiles), a long row and a tail shorter than one tile

Only data is stored; no reference source is copied.
"""
from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from diploma_research_amd import matgen  # noqa: E402
from diploma_research_amd.csr5 import Csr  # noqa: E402
from diploma_research_amd.mmio import read_coo  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    coo = read_coo(os.path.join(HERE, "mm.mtx"))
    csr = Csr.from_coo(coo)
    rows, cols = coo.dimensions.rows, coo.dimensions.cols
    O.ref_h5_store_csr5(os.path.join(HERE, "mm.csr5.h5"), "A", rows, cols, csr.row_start_offsets, csr.col_indices,
                        csr.values)
    O.ref_h5_store_matlab(os.path.join(HERE, "mm.csr.h5"), "A", rows, cols, csr.row_start_offsets, csr.col_indices,
                          csr.values)
    rp, ci, v = matgen.random_csr(900, 700, seed=77, max_len=25, empty_frac=0.15, long_rows=[(40, 600)])
    O.ref_h5_store_csr5(os.path.join(HERE, "dirty.csr5.h5"), "B", 900, 700, rp, ci, v)


if __name__ == "__main__":
    main()
