"""COO -> CSR ingest pinned against the reference's own csr::from_coo
(include/dim/mat/storage_formats/csr.h:42-77, with the symmetric mirroring of
coo.h:38-45), compiled from its header into oracle/_ref: Csr.from_coo and the
CLI's MatrixMarket reader (store-matrix -f csr) must build the same CSR.

The reference sorts each row with std::sort on the column alone
(keyval_iterator.h:56-60, 112-113), which is not stable: the order of
duplicate (row, col) entries among themselves is unspecified there, so for
those only the multiset of values of each duplicate run is compared; all
other arrays and orders are compared bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from diploma_research_amd import h5
from diploma_research_amd.csr5 import Coo, Csr, Dimensions
from diploma_research_amd.mmio import read_coo
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DIM = os.path.join(ROOT, "diploma_research_amd", "dim")

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")


def _coo(rows, cols, nnz, seed, symmetric=False, dup_frac=0.0, integer=False):
    rng = np.random.default_rng(seed)
    r = rng.integers(0, rows, nnz).astype(np.uint32)
    c = rng.integers(0, cols, nnz).astype(np.uint32)
    if symmetric:  # a symmetric MatrixMarket file stores the lower triangle
        r, c = np.maximum(r, c), np.minimum(r, c)
    if not dup_frac:  # drop repeated coordinates, keeping the first
        _, first = np.unique(r.astype(np.int64) * cols + c, return_index=True)
        keep = np.sort(first)
        r, c = r[keep], c[keep]
    else:  # repeat a share of the entries with other values
        k = int(dup_frac * len(r))
        pick = rng.integers(0, len(r), k)
        r = np.concatenate([r, r[pick]])
        c = np.concatenate([c, c[pick]])
    v = rng.integers(-9, 10, len(r)).astype(np.float64) if integer else rng.standard_normal(len(r))
    perm = rng.permutation(len(r))  # COO order is arbitrary
    return r[perm], c[perm], v[perm]


def _assert_same_csr(ours, ref, cols):
    rp, ci, v = ours
    rrp, rci, rv = ref
    np.testing.assert_array_equal(rp, rrp)
    np.testing.assert_array_equal(ci, rci)
    # values: bitwise, except that runs of one duplicated (row, col) compare
    # as sorted multisets
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp.astype(np.int64)))
    key = rows.astype(np.int64) * cols + ci.astype(np.int64)
    starts = np.flatnonzero(np.concatenate([[True], key[1:] != key[:-1]]))
    bounds = np.append(starts, len(key))
    a, b = v.copy(), rv.copy()
    for s, e in zip(bounds[:-1], bounds[1:]):
        if e - s > 1:
            a[s:e] = np.sort(a[s:e])
            b[s:e] = np.sort(b[s:e])
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("case", ["random", "symmetric", "duplicates", "symmetric_duplicates", "empty_rows",
                                  "rectangular"])
def test_from_coo_matches_reference(case):
    rows, cols, nnz, kw = {
        "random": (3000, 3000, 40000, {}),
        "symmetric": (2500, 2500, 30000, {"symmetric": True}),
        "duplicates": (1500, 1500, 30000, {"dup_frac": 0.3}),
        "symmetric_duplicates": (1200, 1200, 20000, {"symmetric": True, "dup_frac": 0.2}),
        "empty_rows": (20000, 20000, 5000, {}),
        "rectangular": (700, 5000, 20000, {}),
    }[case]
    r, c, v = _coo(rows, cols, nnz, seed=len(case), **kw)
    sym = kw.get("symmetric", False)
    ours = Csr.from_coo(Coo(Dimensions(rows, cols), v, r, c, sym))
    ref = O.ref_csr_from_coo(rows, cols, r, c, v, symmetric=sym)
    _assert_same_csr((ours.row_start_offsets, ours.col_indices, ours.values), ref, cols)


def _write_mtx(path, rows, cols, r, c, v, symmetric=False, pattern=False):
    field = "pattern" if pattern else "real"
    kind = "symmetric" if symmetric else "general"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {kind}\n% written by the ingest test\n")
        f.write(f"{rows} {cols} {len(r)}\n")
        for i, j, x in zip(r, c, v):
            f.write(f"{i + 1} {j + 1}\n" if pattern else f"{i + 1} {j + 1} {float(x)!r}\n")


@pytest.mark.parametrize("kind", ["general", "symmetric", "pattern", "duplicates"])
def test_matrix_market_ingest_matches_reference(tmp_path, kind):
    """The Python reader and the CLI's reader (store-matrix -f csr, then the
    MATLAB CSR group read back) against csr::from_coo of the same COO."""
    rows, cols = 900, 1100 if kind == "general" else 900
    r, c, v = _coo(rows, cols, 9000, seed=3, symmetric=kind == "symmetric",
                   dup_frac=0.25 if kind == "duplicates" else 0.0)
    if kind == "pattern":
        v = np.ones(len(v))
    path = tmp_path / f"{kind}.mtx"
    _write_mtx(path, rows, cols, r, c, v, symmetric=kind == "symmetric", pattern=kind == "pattern")
    ref = O.ref_csr_from_coo(rows, cols, r, c, v, symmetric=kind == "symmetric")
    coo = read_coo(path)
    ours = Csr.from_coo(coo)
    _assert_same_csr((ours.row_start_offsets, ours.col_indices, ours.values), ref, cols)
    if not os.path.exists(DIM):
        pytest.skip("dim CLI not built")
    out = tmp_path / "m.csr.h5"
    p = subprocess.run([DIM, "store-matrix", str(path), "-o", str(out), "-f", "csr"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    assert p.returncode == 0, p.stdout
    m = h5.read_csr(out, "A")
    assert (m[0], m[1]) == (rows, cols)
    _assert_same_csr((m[2], m[3], m[4]), ref, cols)
