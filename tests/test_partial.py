"""One rank's share of a CSR5 file: dim_h5_load_csr5_partial against the
reference's own load_csr5_partial (libs/dim/io/h5.cpp:137-173) and its
csr5::spmv<partial> (csr5.h:585-594, 756-782), both compiled from the
reference into oracle/_ref. CPU tests pin the loader (rows, offsets, tail,
the elements) and the meaning of the share; the GPU tests multiply the
loaded share on the MI355X."""
import os

import numpy as np
import pytest

from diploma_research_amd import h5, matgen
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

needs_ref = pytest.mark.skipif(not O.ref_h5_available(), reason="reference HDF5 build (oracle/_ref) absent")


def _case(name):
    if name == "dirty":  # empty rows (dirty tiles) and a long row
        rp, ci, v = matgen.random_csr(3000, 3000, seed=3, max_len=30, empty_frac=0.1, long_rows=[(5, 900)])
        return 3000, 3000, rp, ci, v
    if name == "poisson":
        rp, ci, v = matgen.poisson2d(64)
        return 4096, 4096, rp, ci, v
    if name == "sparse_rect":  # few elements per row, most rows empty
        rp, ci, v = matgen.random_csr(500, 800, seed=5, max_len=3, empty_frac=0.6)
        return 500, 800, rp, ci, v
    if name == "fast_track":  # rows longer than a reference tile: fast-track tiles
        rp, ci, v = matgen.random_csr(400, 5000, seed=9, max_len=8, long_rows=[(10, 300), (200, 2000)])
        return 400, 5000, rp, ci, v
    raise KeyError(name)


def _valid(T, i, count):
    part = -(-T // count)
    return not (part * i > T or (i + 1 < count and part * i == T))


@needs_ref
@pytest.mark.parametrize("name", ["dirty", "poisson", "sparse_rect", "fast_track"])
def test_partial_load_matches_reference(tmp_path, name):
    """Every part of 1, 2, 3, 5, 8 and 13: the same first/last row,
    val_offset, skip_tail, first_tile_uncapped and tile count as the reference's loader; the block
    holds the part's elements in CSR order (the in-memory partition of the same
    CSR); the block's product equals the reference's spmv<partial> of the part
    within the parity gate; parts the reference cannot load (it wraps around
    or asserts) are refused; the parts' products add up to the full product."""
    rows, cols, rp, ci, v = _case(name)
    path = tmp_path / "m.csr5.h5"
    O.ref_h5_store_csr5(path, "A", rows, cols, rp, ci, v)
    x = np.random.default_rng(1).uniform(-1, 1, cols)
    y_full = O.csr_spmv(rp, ci, v, x)
    T = len(v) // 64
    uncapped_seen = set()
    for count in (1, 2, 3, 5, 8, 13):
        total = np.zeros(rows)
        complete = True
        for i in range(count):
            if not _valid(T, i, count):
                complete = False
                with pytest.raises(ValueError, match="beyond the last tile"):
                    h5.load_csr5_partial(path, "A", i, count)
                continue
            ref = O.ref_h5_load_csr5_partial(path, "A", i, count, x=x)
            got = h5.load_csr5_partial(path, "A", i, count)
            for key in ("rows", "cols", "first_row", "last_row", "val_offset", "skip_tail", "tile_count",
                        "first_tile_uncapped"):
                assert got[key] == ref[key], (count, i, key, got[key], ref[key])
            uncapped_seen.add(got["first_tile_uncapped"])
            first, v0, blk = h5.partition_block(rows, cols, rp, ci, v, i, count)
            assert (first, v0) == (got["first_row"], got["val_offset"])
            n, c, brp, bci, bv = got["block"]
            assert (n, c) == (ref["last_row"] - ref["first_row"] + 1, cols)
            for a, b in zip(blk, got["block"]):
                np.testing.assert_array_equal(a, b)
            yb = O.csr_spmv(brp, bci, bv, x)
            d = np.linalg.norm(yb - ref["y_block"])
            assert d <= 1e-13 * max(np.linalg.norm(ref["y_block"]), 1e-300), (count, i, d)
            # the oracle's block product over the same elements
            ob = O.csr_block_spmv(rp, ci, v, got["val_offset"], got["val_offset"] + len(bv), x, first, n)
            np.testing.assert_allclose(ob, yb, rtol=0, atol=1e-13 * max(np.abs(yb).max(), 1.0))
            total[first:first + n] += yb
        if complete:
            np.testing.assert_allclose(total, y_full, rtol=0, atol=1e-12 * max(np.abs(y_full).max(), 1.0))
    if name == "poisson":  # short rows: parts begin both at a row start and inside a row
        assert uncapped_seen == {False, True}


@needs_ref
def test_partial_load_of_reference_written_fixture():
    """The committed reference-written file with dirty tiles (group B): every
    part of 4 equals the reference loader's view."""
    path = os.path.join(GOLDEN, "dirty.csr5.h5")
    whole = h5.read_csr(path, "B")
    x = np.random.default_rng(2).uniform(-1, 1, whole[1])
    for i in range(4):
        ref = O.ref_h5_load_csr5_partial(path, "B", i, 4, x=x)
        got = h5.load_csr5_partial(path, "B", i, 4)
        assert (got["first_row"], got["last_row"], got["val_offset"]) == (ref["first_row"], ref["last_row"],
                                                                           ref["val_offset"])
        n, c, brp, bci, bv = got["block"]
        yb = O.csr_spmv(brp, bci, bv, x)
        np.testing.assert_allclose(yb, ref["y_block"], rtol=0, atol=1e-13 * max(np.abs(yb).max(), 1.0))


def test_partial_load_rejects_bad_arguments(tmp_path):
    rows, cols, rp, ci, v = _case("poisson")
    path = tmp_path / "m.h5"
    h5.write_csr5(path, "A", rows, cols, rp, ci, v)
    with pytest.raises(ValueError, match="part_idx"):
        h5.load_csr5_partial(path, "A", 3, 3)
    h5.write_matlab_csr(path, "B", rows, cols, rp, ci, v, append=True)
    with pytest.raises(ValueError, match="no CSR5"):
        h5.load_csr5_partial(path, "B", 0, 2)


# ------------------------------------------------------------------ GPU ---

@needs_ref
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dirty", "poisson", "fast_track"])
@pytest.mark.parametrize("sigma", [4, 7, 16])
def test_partial_share_on_gpu_matches_reference(cuda, tmp_path, name, sigma):
    """Each loaded part of 5, converted and multiplied on the GPU (as one
    rank's matrix), against the reference's spmv<partial> of that part."""
    from gpu_util import assert_spmv_close, to_dev
    from diploma_research_amd import Csr5

    rows, cols, rp, ci, v = _case(name)
    path = tmp_path / "m.csr5.h5"
    O.ref_h5_store_csr5(path, "A", rows, cols, rp, ci, v)
    x = np.random.default_rng(sigma).uniform(-1, 1, cols)
    T = len(v) // 64
    for i in range(5):
        if not _valid(T, i, 5):
            continue
        ref = O.ref_h5_load_csr5_partial(path, "A", i, 5, x=x)
        n, c, brp, bci, bv = h5.load_csr5_partial(path, "A", i, 5)["block"]
        A = Csr5.from_csr((n, c, brp, bci, bv), sigma=sigma)
        y = A.spmv(to_dev(x)).cpu().numpy()
        assert_spmv_close(y, ref["y_block"], brp, bci, bv, x)


@needs_ref
@pytest.mark.gpu
def test_load_h5_to_gpu_in_one_call(cuda, tmp_path):
    """dim_csr5_load_h5 and dim_csr5_load_h5_partial: the whole matrix and
    each part of 3 loaded straight to GPU matrices, against the reference's
    load_csr5 product and spmv<partial> of each part."""
    from gpu_util import assert_spmv_close, to_dev

    rows, cols, rp, ci, v = _case("dirty")
    path = tmp_path / "m.csr5.h5"
    O.ref_h5_store_csr5(path, "A", rows, cols, rp, ci, v)
    x = np.random.default_rng(4).uniform(-1, 1, cols)
    A = h5.load_csr5_gpu(path, "A", sigma=8)
    assert (A.dimensions.rows, A.dimensions.cols, A.non_zero_count()) == (rows, cols, len(v))
    assert_spmv_close(A.spmv(to_dev(x)).cpu().numpy(), O.csr_spmv(rp, ci, v, x), rp, ci, v, x)
    for i in range(3):
        ref = O.ref_h5_load_csr5_partial(path, "A", i, 3, x=x)
        first, P = h5.load_csr5_partial_gpu(path, "A", i, 3, sigma=8)
        assert first == ref["first_row"] and P.dimensions.rows == ref["last_row"] - ref["first_row"] + 1
        n, c, brp, bci, bv = h5.load_csr5_partial(path, "A", i, 3)["block"]
        assert_spmv_close(P.spmv(to_dev(x)).cpu().numpy(), ref["y_block"], brp, bci, bv, x)


@needs_ref
@pytest.mark.gpu
def test_cli_spmv_of_one_part(cuda, tmp_path):
    """`dim spmv m.csr5.h5 --part i/n` multiplies one rank's share (the
    partial strategy) and writes its rows, equal to the reference's
    spmv<partial> of that part."""
    import subprocess

    from gpu_util import assert_spmv_close

    dim = os.path.join(ROOT, "diploma_research_amd", "dim")
    rows, cols, rp, ci, v = _case("dirty")
    path = tmp_path / "m.csr5.h5"
    O.ref_h5_store_csr5(path, "A", rows, cols, rp, ci, v)
    x = np.random.default_rng(6).uniform(-1, 1, cols)
    h5.write_vector(path, "X", x)
    for i in range(3):
        out = tmp_path / f"y{i}.h5"
        p = subprocess.run([dim, "spmv", str(path), "--part", f"{i}/3", "-n", "3", "--x-dataset", "X",
                            "--output-file", str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=300)
        assert p.returncode == 0, p.stdout
        ref = O.ref_h5_load_csr5_partial(path, "A", i, 3, x=x)
        assert f"rows {ref['first_row']}..{ref['last_row']}" in p.stdout, p.stdout
        n, c, brp, bci, bv = h5.load_csr5_partial(path, "A", i, 3)["block"]
        assert_spmv_close(h5.read_vector(out, "Y"), ref["y_block"], brp, bci, bv, x)
