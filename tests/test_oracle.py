"""The oracle pinned against the reference: its own known-answer tests
(tests/test_csr5.cpp, tests/test_simd.cpp), the golden fixtures written from
the compiled reference header, brute force, and closed forms."""
import numpy as np
import pytest

from conftest import load_golden
from diploma_research_amd import matgen
from diploma_research_amd.csr5 import Csr
from diploma_research_amd.mmio import read_coo
from oracle import oracle as O

GOLDEN_SMALL = ["mm", "rand_dirty", "rand_longrows", "rect_exact"]


def _oracle_from(g, sigma=16, omega=4):
    return O.Csr5(int(g["rows"]), int(g["cols"]), g["row_ptr"], g["col_idx"], g["vals"], sigma, omega)


# ----------------------------------------------------- reference KATs ---

def test_mm_structure_kat(golden_dir):
    """tests/test_csr5.cpp:26-58 on tests/test_files/mm.mtx."""
    coo = read_coo(f"{golden_dir}/mm.mtx")
    csr = Csr.from_coo(coo)
    A = O.Csr5(coo.dimensions.rows, coo.dimensions.cols, csr.row_start_offsets, csr.col_indices, csr.values)
    info = A.info
    assert info["tile_count"] == 2
    assert (info["tile_ptr"][-1] & O.IDX_MASK) == 154  # tail_partition_start
    assert not np.array_equal(info["vals"], csr.values)  # transposed
    bf = info["bit_flag"]
    assert list(bf[0]) == [0b1000_0100_0010_0001, 0b1000_1000_1000_1000, 0b1010_1000_1000_1000,
                           0b1010_1010_1010_1010]
    assert list(info["y_offset"][0]) == [0, 4, 8, 13]
    assert list(info["scansum_offset"][0]) == [0, 0, 0, 0]
    assert list(bf[1]) == [0b1110, 0, 0, 0]
    assert list(info["y_offset"][1]) == [0, 4, 4, 4]
    assert list(info["scansum_offset"][1]) == [3, 0, 0, 0]
    assert len(info["tile_desc_offset_ptr"]) == info["tile_count"] + 1
    assert len(info["tile_desc_offset"]) == info["y_offset"][1][3] + 1
    assert (info["tile_ptr"][1] & O.IDX_MASK) != 0
    assert info["tile_desc_offset_ptr"][1] == 0
    assert info["tile_desc_offset"][3] == 154 - (info["tile_ptr"][1] & O.IDX_MASK)


def test_num_bits_set_kat():
    """tests/test_csr5.cpp:15-24: popcount with the implicit first bit."""
    # a 1-tile matrix whose column 1 has bits 0b1000100010001000 and column 0 starts with a row
    rows = 5
    # element starts of rows inside a 64-element tile: column c covers [16c, 16c+16)
    starts = [0, 16 + 3, 16 + 7, 16 + 11, 16 + 15]
    rp = np.array(starts + [64], np.uint32)
    A = O.Csr5(rows, 1, rp, np.zeros(64, np.uint32), np.ones(64))
    info = A.info
    assert info["bit_flag"][0][1] == 0b1000_1000_1000_1000
    assert info["y_offset"][0][1] == 1  # column 0 counts its implicit bit


def test_simd_kats():
    """tests/test_simd.cpp:17-30."""
    L = O.lib()
    a = np.array([1.0, 2.0, 3.0, 4.0])
    s = np.zeros(4)
    L.orc_lane_inclusive_scan(4, a, s)
    assert list(s) == [1.0, 3.0, 6.0, 10.0]
    out = np.zeros(4)
    L.orc_shuffle_relative(4, a, np.array([0, 0, 0, 0], np.uint32), out)
    assert list(out) == [1, 2, 3, 4]
    L.orc_shuffle_relative(4, a, np.array([1, 1, 1, 0], np.uint32), out)
    assert list(out) == [2, 3, 4, 4]
    L.orc_shuffle_relative(4, a, np.array([3, 2, 1, 0], np.uint32), out)
    assert list(out) == [4, 4, 4, 4]
    assert L.orc_lane_sum(4, a) == 10.0


def test_fast_segmented_sum_closed_form():
    """out[i] = sum of vec[i+1 .. i+1+off[i]] (vec shifted left, last lane 0)."""
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(50):
        v = rng.integers(-9, 10, 4).astype(np.float64)
        off = np.array([rng.integers(0, 4 - i) for i in range(4)], np.uint32)
        out = np.zeros(4)
        L.orc_fast_segmented_sum(4, v, off, out)
        t = np.append(v[1:], 0.0)
        exp = [t[i: i + off[i] + 1].sum() for i in range(4)]
        assert list(out) == exp


# --------------------------------------- golden fixtures (reference) ---

@pytest.mark.parametrize("name", GOLDEN_SMALL)
def test_conversion_matches_reference_fixture(name):
    g = load_golden(name)
    A = _oracle_from(g)
    info = A.info
    assert info["tile_count"] == int(g["tile_count"])
    np.testing.assert_array_equal(info["tile_ptr"], g["tile_ptr"])
    if info["tile_count"]:
        np.testing.assert_array_equal(A.packed_desc(), g["tile_desc"])
    np.testing.assert_array_equal(info["tile_desc_offset_ptr"], g["tile_desc_offset_ptr"])
    np.testing.assert_array_equal(info["tile_desc_offset"], g["tile_desc_offset"])
    np.testing.assert_array_equal(info["vals"], g["t_vals"])
    np.testing.assert_array_equal(info["col_idx"], g["t_col_idx"])


@pytest.mark.parametrize("name", GOLDEN_SMALL)
def test_spmv_bit_identical_to_reference_fixture(name):
    g = load_golden(name)
    A = _oracle_from(g)
    np.testing.assert_array_equal(A.spmv(g["x"], nthreads=1), g["y_ref_t1"])
    np.testing.assert_array_equal(A.spmv(g["x"], nthreads=8), g["y_ref_t8"])


def test_poisson2d_fixture():
    g = load_golden("poisson2d_256")
    rp, ci, v = matgen.poisson2d(256)
    A = O.Csr5(65536, 65536, rp, ci, v)
    import hashlib
    info = A.info
    assert hashlib.sha256(info["tile_ptr"].tobytes()).hexdigest() == str(g["tile_ptr_sha"])
    assert hashlib.sha256(A.packed_desc().tobytes()).hexdigest() == str(g["tile_desc_sha"])
    x = np.random.default_rng(int(g["x_seed"])).uniform(-1.0, 1.0, 65536)
    np.testing.assert_array_equal(A.spmv(x, nthreads=8), g["y_ref_t8"])
    np.testing.assert_array_equal(A.spmv(x, nthreads=1), g["y_ref_t1"])


# ---------------------------------------------- live reference (_ref) ---

@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("seed", range(6))
def test_oracle_bit_identical_to_reference_random(seed):
    rng = np.random.default_rng(seed)
    rows = int(rng.integers(50, 3000))
    cols = int(rng.integers(50, 3000))
    rp, ci, v = matgen.random_csr(rows, cols, seed, max_len=int(rng.integers(1, 40)),
                                  empty_frac=float(rng.choice([0, 0.05, 0.4])),
                                  long_rows=[(int(rng.integers(rows)), int(rng.integers(64, 2000)))])
    A = O.Csr5(rows, cols, rp, ci, v)
    R = O.RefCsr5(rows, cols, rp, ci, v)
    ia, ir = A.info, R.info
    np.testing.assert_array_equal(ia["tile_ptr"], ir["tile_ptr"])
    if ia["tile_count"]:
        np.testing.assert_array_equal(A.packed_desc(), ir["tile_desc"])
    np.testing.assert_array_equal(ia["tile_desc_offset"], ir["tile_desc_offset"])
    np.testing.assert_array_equal(ia["vals"], ir["vals"])
    x = rng.uniform(-1, 1, cols)
    for th in (1, 3, 8):
        np.testing.assert_array_equal(A.spmv(x, th), R.spmv(x, th))


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_reference_stale_state_defect_documented():
    """The reference's own calling convention (y and calibrator reused across
    calls, benchmarks/csr5/dim/main.cpp:197-209) returns wrong rows from the
    second call on; the oracle resets both on every call."""
    rp, ci, v = matgen.random_csr(1000, 1000, 1, max_len=8)
    R = O.RefCsr5(1000, 1000, rp, ci, v)
    x = np.random.default_rng(2).uniform(-1, 1, 1000)
    exact = O.csr_spmv(rp, ci, v, x)
    s = R.session(4)
    y = np.zeros(1000)
    s.spmv(x, y)
    s.spmv(x, y)
    assert np.abs(y - exact).max() > 1e-3  # stale: wrong
    A = O.Csr5(1000, 1000, rp, ci, v)
    A.spmv(x, 4)
    np.testing.assert_allclose(A.spmv(x, 4), exact, rtol=0, atol=1e-13)


# --------------------------------- the original CSR5 (second oracle) ---

@pytest.mark.skipif(not O.orig_available(), reason="original CSR5 build absent")
@pytest.mark.parametrize("case", range(5))
def test_original_csr5_integer_exact(case):
    """The original CSR5 the reference vendors (anonymouslib, omega = 4, its
    compiled-in sigma = 16) agrees exactly with the plain CSR product and
    with the oracle's restatement of the reference on integer inputs (empty
    rows, long rows, a tail shorter than a tile, tiny matrices)."""
    rows, cols, ml, ef, lr = [(2000, 1500, 12, 0.1, [(7, 1400), (8, 300)]), (333, 777, 3, 0.5, []), (5, 5, 2, 0.0, []),
                              (1, 300, 0, 0.0, [(0, 300)]), (4096, 64, 40, 0.02, [(0, 64), (4095, 64)])][case]
    rp, ci, v = matgen.random_csr(rows, cols, seed=50 + case, max_len=ml, empty_frac=ef, long_rows=lr, integer=True)
    x = np.random.default_rng(case).integers(-5, 6, cols).astype(np.float64)
    exact = O.csr_spmv(rp, ci, v, x)
    np.testing.assert_array_equal(O.orig_spmv(rows, cols, rp, ci, v, x), exact)
    np.testing.assert_array_equal(O.Csr5(rows, cols, rp, ci, v, 16, 4).spmv(x, 4), exact)
    with pytest.raises(ValueError, match="sigma = 16"):
        O.orig_spmv(rows, cols, rp, ci, v, x, 8)


@pytest.mark.skipif(not O.orig_available(), reason="original CSR5 build absent")
@pytest.mark.parametrize("seed", range(4))
def test_original_csr5_agrees_with_oracle_random(seed):
    """Floating-point inputs: the original and the oracle sum each row in
    their own orders; both stay within the componentwise rounding bound
    (len_i + 2) * u * sum |a x| of the exact product, and within 1e-13
    normwise of each other."""
    rng = np.random.default_rng(seed)
    rows, cols = int(rng.integers(100, 4000)), int(rng.integers(100, 4000))
    rp, ci, v = matgen.random_csr(rows, cols, seed, max_len=int(rng.integers(1, 40)),
                                  empty_frac=float(rng.choice([0, 0.05, 0.4])),
                                  long_rows=[(int(rng.integers(rows)), int(rng.integers(64, min(cols, 3000))))])
    x = rng.uniform(-1, 1, cols)
    yo = O.orig_spmv(rows, cols, rp, ci, v, x, 16)
    ya = O.Csr5(rows, cols, rp, ci, v).spmv(x, 4)
    lens = np.diff(rp.astype(np.int64))
    absum = np.bincount(np.repeat(np.arange(rows), lens), weights=np.abs(v * x[ci]), minlength=rows)
    exact = O.csr_spmv(rp, ci, v, x)
    u = np.finfo(np.float64).eps / 2
    for y in (yo, ya):
        assert (np.abs(y - exact) <= (lens + 2) * u * absum + 1e-300).all()
    assert np.linalg.norm(yo - ya) <= 1e-13 * np.linalg.norm(exact)


# ------------------------------------------------ brute force / closed ---

@pytest.mark.parametrize("sigma,omega", [(16, 4), (8, 4), (16, 64), (32, 64), (4, 8)])
def test_oracle_integer_exact_any_tiling(sigma, omega):
    """Integer-valued A and x make every summation order exact."""
    rp, ci, v = matgen.random_csr(2000, 1500, seed=sigma + omega, max_len=12, empty_frac=0.1,
                                  long_rows=[(7, 1400), (8, 300)], integer=True)
    x = np.random.default_rng(9).integers(-5, 6, 1500).astype(np.float64)
    A = O.Csr5(2000, 1500, rp, ci, v, sigma, omega)
    exact = O.csr_spmv(rp, ci, v, x)
    np.testing.assert_array_equal(A.spmv(x, 1), exact)
    np.testing.assert_array_equal(A.spmv(x, 7), exact)
    dense = np.zeros((2000, 1500))
    dense[np.repeat(np.arange(2000), np.diff(rp.astype(np.int64))), ci] = v
    np.testing.assert_array_equal(dense @ x, exact)


def test_poisson_closed_forms():
    """A*1 = number of missing neighbours (boundary count); A*v = lambda*v for
    the discrete sine mode."""
    side = 40
    rp, ci, v = matgen.poisson2d(side)
    A = O.Csr5(side * side, side * side, rp, ci, v)
    y = A.spmv(np.ones(side * side))
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    missing = (i == 0).astype(int) + (i == side - 1) + (j == 0) + (j == side - 1)
    np.testing.assert_array_equal(y, missing.ravel().astype(float))
    p, q = 3, 5
    h = np.pi / (side + 1)
    mode = (np.sin(p * h * (i + 1)) * np.sin(q * h * (j + 1))).ravel()
    lam = 4 - 2 * np.cos(p * h) - 2 * np.cos(q * h)
    np.testing.assert_allclose(A.spmv(mode), lam * mode, atol=1e-13)


def test_size_mismatch_raises():
    rp, ci, v = matgen.random_csr(100, 80, 0)
    A = O.Csr5(100, 80, rp, ci, v)
    with pytest.raises(ValueError):
        A.spmv(np.zeros(81))


def test_block_oracle_sums_to_full():
    """Row-block shares over a tile partition add up to the full product."""
    rp, ci, v = matgen.random_csr(3000, 3000, 4, max_len=20, empty_frac=0.1, long_rows=[(100, 2500)])
    x = np.random.default_rng(1).uniform(-1, 1, 3000)
    full = O.csr_spmv(rp, ci, v, x)
    nnz = len(v)
    acc = np.zeros(3000)
    bounds = [0, nnz // 5, nnz // 2, (3 * nnz) // 4, nnz]
    for b, e in zip(bounds[:-1], bounds[1:]):
        first = int(np.searchsorted(rp, b, side="right") - 1)
        last = int(np.searchsorted(rp, max(e - 1, b), side="right") - 1)
        yb = O.csr_block_spmv(rp, ci, v, b, e, x, first, last - first + 1)
        acc[first: last + 1] += yb
    np.testing.assert_allclose(acc, full, atol=1e-12)


# ------------------------------------------------------------------ CG ---

def test_cg_diagonal_distinct_eigenvalues():
    """CG terminates in (number of distinct eigenvalues) steps in exact arithmetic."""
    n = 500
    d = np.repeat([1.0, 2.0, 5.0, 11.0], n // 4)
    rp = np.arange(n + 1, dtype=np.uint32)
    A = O.Csr5(n, n, rp, np.arange(n, dtype=np.uint32), d)
    b = np.random.default_rng(0).uniform(-1, 1, n)
    r = A.cg(b, rtol=1e-12, maxit=50)
    assert r["converged"] and r["iterations"] == 4
    np.testing.assert_allclose(r["x"], b / d, rtol=1e-12)


def test_cg_iterates_minimise_energy_over_krylov_space():
    """x_k = argmin ||x - x*||_A over K_k(A, b) (x0 = 0): checked against a
    dense projection onto an orthonormal Krylov basis."""
    side = 12
    rp, ci, v = matgen.poisson2d(side)
    n = side * side
    A = O.Csr5(n, n, rp, ci, v)
    dense = np.zeros((n, n))
    dense[np.repeat(np.arange(n), np.diff(rp.astype(np.int64))), ci] = v
    b = np.random.default_rng(5).uniform(-1, 1, n)
    for k in (1, 3, 7):
        r = A.cg(b, rtol=0.0, maxit=k)
        assert r["iterations"] == k
        K = np.empty((n, k))
        w = b.copy()
        for i in range(k):
            K[:, i] = w
            w = dense @ w
        Q, _ = np.linalg.qr(K)
        y = np.linalg.solve(Q.T @ dense @ Q, Q.T @ b)
        np.testing.assert_allclose(r["x"], Q @ y, rtol=1e-8, atol=1e-10)


def test_cg_poisson_converges_with_true_residual():
    side = 48
    rp, ci, v = matgen.poisson2d(side)
    n = side * side
    A = O.Csr5(n, n, rp, ci, v)
    xs = np.random.default_rng(1).uniform(-1, 1, n)
    b = A.spmv(xs)
    r = A.cg(b, rtol=1e-10, maxit=2000, history=True)
    assert r["converged"]
    assert r["relres"] <= 1e-10 and r["relres_true"] <= 2e-10
    assert r["history"][0] == pytest.approx(1.0)
    np.testing.assert_allclose(r["x"], xs, atol=1e-7)
