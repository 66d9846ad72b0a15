"""HDF5 matrix files (libdim_h5, include/dim_h5.h) against the reference's
own HDF5 I/O: fixtures written by the reference (tests/golden/*.h5, from
tests/golden/make_h5_golden.py) are read back to the CSR they were made from,
and files written here are loaded by the reference's h5::load_csr5 /
read_matlab_compatible (oracle/_ref/libdimref_h5.so) with arrays identical to
the ones its own writer stores."""
import os

import numpy as np
import pytest

from diploma_research_amd import h5, matgen
from diploma_research_amd.csr5 import Csr
from diploma_research_amd.mmio import read_coo
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not O.ref_h5_available(), reason="oracle/_ref/libdimref_h5.so not built")


def _mm():
    coo = read_coo(os.path.join(GOLDEN, "mm.mtx"))
    c = Csr.from_coo(coo)
    return coo.dimensions.rows, coo.dimensions.cols, c.row_start_offsets, c.col_indices, c.values


def _dirty():
    rp, ci, v = matgen.random_csr(900, 700, seed=77, max_len=25, empty_frac=0.15, long_rows=[(40, 600)])
    return 900, 700, rp, ci, v


def _assert_csr(got, want):
    assert got[0] == want[0] and got[1] == want[1]
    np.testing.assert_array_equal(got[2], np.asarray(want[2], np.uint32))
    np.testing.assert_array_equal(got[3], np.asarray(want[3], np.uint32))
    np.testing.assert_array_equal(got[4], np.asarray(want[4], np.float64))


@pytest.mark.parametrize("name,group,make", [("mm.csr5.h5", "A", _mm), ("mm.csr.h5", "A", _mm),
                                             ("dirty.csr5.h5", "B", _dirty)])
def test_reads_reference_written_fixtures(name, group, make):
    path = os.path.join(GOLDEN, name)
    assert h5.is_hdf5(path)
    assert h5.group_kind(path, group) == (h5.CSR5 if "csr5" in name else h5.MATLAB_CSR)
    _assert_csr(h5.read_csr(path, group), make())


@needs_ref
@pytest.mark.parametrize("make", [_mm, _dirty])
def test_csr5_group_identical_to_reference_writer(tmp_path, make):
    rows, cols, rp, ci, v = make()
    ours, theirs = tmp_path / "ours.h5", tmp_path / "ref.h5"
    h5.write_csr5(ours, "A", rows, cols, rp, ci, v)
    O.ref_h5_store_csr5(theirs, "A", rows, cols, rp, ci, v)
    a, b = O.ref_h5_load_csr5(ours), O.ref_h5_load_csr5(theirs)
    for k in b:
        if isinstance(b[k], np.ndarray):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        else:
            assert a[k] == b[k], k


@needs_ref
def test_csr5_group_random_shapes(tmp_path):
    """Fast-track tiles (one long row over whole tiles), dirty tiles, tails."""
    rng = np.random.default_rng(3)
    for i in range(6):
        rows, cols = int(rng.integers(1, 400)), int(rng.integers(1, 300))
        rp, ci, v = matgen.random_csr(rows, cols, seed=100 + i, max_len=int(rng.integers(1, 50)),
                                      empty_frac=float(rng.uniform(0, 0.4)),
                                      long_rows=[(int(rng.integers(0, rows)), int(rng.integers(64, 300)))])
        ours, theirs = tmp_path / f"o{i}.h5", tmp_path / f"r{i}.h5"
        h5.write_csr5(ours, "A", rows, cols, rp, ci, v)
        O.ref_h5_store_csr5(theirs, "A", rows, cols, rp, ci, v)
        a, b = O.ref_h5_load_csr5(ours), O.ref_h5_load_csr5(theirs)
        for k in b:
            if isinstance(b[k], np.ndarray):
                np.testing.assert_array_equal(a[k], b[k], err_msg=f"case {i}: {k}")
        _assert_csr(h5.read_csr(theirs, "A"), (rows, cols, rp, ci, v))


@needs_ref
def test_matlab_group_both_ways(tmp_path):
    rows, cols, rp, ci, v = _dirty()
    h5.write_matlab_csr(tmp_path / "m.h5", "A", rows, cols, rp, ci, v)
    _assert_csr(O.ref_h5_read_matlab(tmp_path / "m.h5", "A"), (rows, cols, rp, ci, v))
    O.ref_h5_store_matlab(tmp_path / "r.h5", "A", rows, cols, rp, ci, v)
    _assert_csr(h5.read_csr(tmp_path / "r.h5", "A"), (rows, cols, rp, ci, v))


def test_mm_tile_descriptors_match_known_answers(tmp_path):
    """tests/test_csr5.cpp:40-57: tile 0 columns {y0,s0,0x8421} {y4,s0,0x8888}
    {y8,s0,0xa888} {y13,s0,0xaaaa}; tile 1 column 0 {y0,s3,0x000e}; tile_ptr[1]
    = 20 and dirty (read from the stored words of our own writer)."""
    if not O.ref_h5_available():
        pytest.skip("reference loader not built")
    rows, cols, rp, ci, v = _mm()
    h5.write_csr5(tmp_path / "mm.h5", "A", rows, cols, rp, ci, v)
    d = O.ref_h5_load_csr5(tmp_path / "mm.h5")
    w = d["tile_desc"]

    def col(t, c):
        x = int(w[4 * t + c])
        return x & 0x7F, (x >> 7) & 0x7, (x >> 10) & 0xFFFF

    assert [col(0, c) for c in range(4)] == [(0, 0, 0x8421), (4, 0, 0x8888), (8, 0, 0xA888), (13, 0, 0xAAAA)]
    assert col(1, 0) == (0, 3, 0x000E)
    assert d["tile_ptr"][1] & O.IDX_MASK == 20 and d["tile_ptr"][1] & O.DIRTY_BIT


def test_vectors_round_trip_and_overwrite(tmp_path):
    p = tmp_path / "v.h5"
    x = np.random.default_rng(0).standard_normal(1001)
    h5.write_vector(p, "Y", x)
    np.testing.assert_array_equal(h5.read_vector(p, "Y"), x)
    with pytest.raises(h5.H5Error, match="overwrite"):
        h5.write_vector(p, "Y", x + 1)
    h5.write_vector(p, "Y", x + 1, overwrite=True)
    np.testing.assert_array_equal(h5.read_vector(p, "/Y"), x + 1)
    h5.write_vector(p, "grp/sub/Z", x[:5])
    np.testing.assert_array_equal(h5.read_vector(p, "grp/sub/Z"), x[:5])
    h5.write_vector(p, "E", np.zeros(0))
    assert h5.read_vector(p, "E").size == 0


def test_chunked_and_compressed_groups_read_back(tmp_path):
    rows, cols, rp, ci, v = _dirty()
    h5.write_csr5(tmp_path / "c.h5", "A", rows, cols, rp, ci, v, chunk={"vals": 1000, "col_idx": 1000},
                  deflate={"vals": 4, "tile_desc": 1})
    _assert_csr(h5.read_csr(tmp_path / "c.h5", "A"), (rows, cols, rp, ci, v))
    h5.write_matlab_csr(tmp_path / "c.h5", "M", rows, cols, rp, ci, v, append=True, chunk={"values": 500})
    _assert_csr(h5.read_csr(tmp_path / "c.h5", "M"), (rows, cols, rp, ci, v))
    _assert_csr(h5.read_csr(tmp_path / "c.h5", "A"), (rows, cols, rp, ci, v))  # append kept group A


def test_errors(tmp_path):
    with pytest.raises(h5.H5Error):
        h5.read_csr(tmp_path / "missing.h5", "A")
    rows, cols, rp, ci, v = _dirty()
    h5.write_csr5(tmp_path / "e.h5", "A", rows, cols, rp, ci, v)
    with pytest.raises(h5.H5Error, match="no group"):
        h5.read_csr(tmp_path / "e.h5", "Q")
    with pytest.raises(ValueError, match="exists"):
        h5.write_csr5(tmp_path / "e.h5", "A", rows, cols, rp, ci, v, append=True)
    bad = np.asarray(ci, np.uint32).copy()
    bad[0] = cols
    with pytest.raises(ValueError, match="out of range"):
        h5.write_csr5(tmp_path / "f.h5", "A", rows, cols, rp, bad, v)
    assert not h5.is_hdf5(os.path.join(GOLDEN, "mm.mtx"))


def test_library_exports_every_declared_symbol():
    import re

    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "include", "dim_h5.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"\b(dim_h5_\w+|dim_csr5_load_h5\w*)\s*\(", hdr))
    assert declared == set(h5.EXPORTS)
    L = h5.lib()
    for name in declared:
        assert hasattr(L, name), name
