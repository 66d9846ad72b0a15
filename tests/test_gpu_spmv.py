"""GPU parity of the CSR5 product (HIP kernels via the C ABI) against the
oracle and the reference's golden fixtures."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden
from gpu_util import assert_spmv_close, to_dev
from diploma_research_amd import Csr5, matgen
from diploma_research_amd import native as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SIGMAS = [4, 7, 8, 16, 24, 32]


@pytest.mark.parametrize("name", ["mm", "rand_dirty", "rand_longrows", "rect_exact"])
@pytest.mark.parametrize("sigma", SIGMAS)
def test_golden_fixtures(cuda, name, sigma):
    g = load_golden(name)
    rows, cols = int(g["rows"]), int(g["cols"])
    A = Csr5.from_csr((rows, cols, g["row_ptr"], g["col_idx"], g["vals"]), sigma=sigma)
    y = A.spmv(to_dev(g["x"])).cpu().numpy()
    assert_spmv_close(y, g["y_ref_t8"], g["row_ptr"], g["col_idx"], g["vals"], g["x"])


@pytest.mark.parametrize("sigma", SIGMAS)
def test_poisson2d_fixture(cuda, sigma):
    g = load_golden("poisson2d_256")
    rp, ci, v = matgen.poisson2d(256)
    x = np.random.default_rng(int(g["x_seed"])).uniform(-1.0, 1.0, 65536)
    A = Csr5.from_csr((65536, 65536, rp, ci, v), sigma=sigma)
    y = A.spmv(to_dev(x)).cpu().numpy()
    assert_spmv_close(y, g["y_ref_t8"], rp, ci, v, x)


CASES = [
    dict(rows=1000, cols=1000, max_len=9, empty_frac=0.15, long_rows=[(3, 300), (500, 700)]),
    dict(rows=300, cols=5000, max_len=4, empty_frac=0.0, long_rows=[(10, 4000), (11, 64), (200, 1000)]),
    dict(rows=5000, cols=300, max_len=3, empty_frac=0.6, long_rows=[]),
    dict(rows=1, cols=20000, max_len=0, empty_frac=0.0, long_rows=[(0, 20000)]),
    dict(rows=20000, cols=1, max_len=1, empty_frac=0.0, long_rows=[]),
    dict(rows=4096, cols=4096, max_len=64, empty_frac=0.02, long_rows=[(0, 4096), (4095, 4096)]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("sigma", SIGMAS)
def test_integer_inputs_bit_exact(cuda, case, sigma):
    """Integer-valued A and x: every order of summation is exact, so the GPU
    result must equal the oracle bit for bit (empty rows, long rows, dirty and
    fast-track tiles, padded last tile, rectangular shapes)."""
    c = CASES[case]
    rp, ci, v = matgen.random_csr(c["rows"], c["cols"], seed=100 + case, max_len=c["max_len"],
                                  empty_frac=c["empty_frac"], long_rows=c["long_rows"], integer=True)
    x = np.random.default_rng(case).integers(-7, 8, c["cols"]).astype(np.float64)
    A = Csr5.from_csr((c["rows"], c["cols"], rp, ci, v), sigma=sigma)
    yg = A.spmv(to_dev(x)).cpu().numpy()
    ref = O.Csr5(c["rows"], c["cols"], rp, ci, v).spmv(x, nthreads=4)
    np.testing.assert_array_equal(yg, ref)


@pytest.mark.parametrize("sigma", SIGMAS)
def test_format_matches_oracle_conversion(cuda, sigma):
    """The GPU format's tile pointers, bit flags and segment->row maps equal
    the oracle's conversion at omega=64 on every full tile."""
    rp, ci, v = matgen.random_csr(3000, 3000, seed=sigma, max_len=20, empty_frac=0.1, long_rows=[(5, 2900)])
    A = Csr5.from_csr((3000, 3000, rp, ci, v), sigma=sigma)
    f = A.format_arrays()
    o = O.Csr5(3000, 3000, rp, ci, v, sigma, 64).info
    T = o["tile_count"]  # full tiles
    np.testing.assert_array_equal(f["tile_ptr"][:T], o["tile_ptr"][:T])
    np.testing.assert_array_equal(f["bit_flag"][:T], o["bit_flag"][:T].astype(np.uint32))
    # segment k of a dirty tile maps to the same relative row in both formats
    for t in range(T):
        if o["tile_ptr"][t] & O.DIRTY_BIT:
            a = o["tile_desc_offset"][o["tile_desc_offset_ptr"][t]:]
            b = f["tile_desc_offset"][f["tile_desc_offset_ptr"][t]:f["tile_desc_offset_ptr"][t + 1]]
            np.testing.assert_array_equal(b, a[: len(b)])
    np.testing.assert_array_equal(f["empty_rows"], np.nonzero(np.diff(rp.astype(np.int64)) == 0)[0])


@pytest.mark.skipif(not O.orig_available(), reason="original CSR5 build absent")
@pytest.mark.parametrize("sigma", [7, 16, 32])
def test_gpu_matches_original_csr5(cuda, sigma):
    """Against the second CPU oracle, the original CSR5 that the reference
    vendors (omega = 4, sigma = 16): normwise within the parity gate."""
    rp, ci, v = matgen.random_csr(6000, 5000, seed=sigma, max_len=30, empty_frac=0.1, long_rows=[(3, 4000)])
    x = np.random.default_rng(sigma).uniform(-1, 1, 5000)
    A = Csr5.from_csr((6000, 5000, rp, ci, v), sigma=sigma)
    assert_spmv_close(A.spmv(to_dev(x)).cpu().numpy(), O.orig_spmv(6000, 5000, rp, ci, v, x), rp, ci, v, x)


def test_overwrites_y_and_is_deterministic(cuda):
    import torch

    rp, ci, v = matgen.random_csr(5000, 5000, seed=3, max_len=30, empty_frac=0.2, long_rows=[(9, 4000)])
    x = np.random.default_rng(0).uniform(-1, 1, 5000)
    A = Csr5.from_csr((5000, 5000, rp, ci, v), sigma=16)
    xd = to_dev(x)
    y1 = torch.full((5000,), float("nan"), dtype=torch.float64, device="cuda")
    A.spmv(xd, y1)
    y2 = torch.full((5000,), 123.0, dtype=torch.float64, device="cuda")
    for _ in range(3):
        A.spmv(xd, y2)
    assert torch.equal(y1, y2)
    assert_spmv_close(y1.cpu().numpy(), O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


def test_size_mismatch_raises_invalid_argument(cuda):
    import torch

    rp, ci, v = matgen.random_csr(100, 80, 0)
    A = Csr5.from_csr((100, 80, rp, ci, v))
    with pytest.raises(ValueError, match="column count"):
        A.spmv(torch.zeros(81, dtype=torch.float64, device="cuda"))


def test_empty_matrix_and_all_empty_rows(cuda):
    import torch

    A = Csr5.from_csr((7, 5, np.zeros(8, np.uint32), np.zeros(0, np.uint32), np.zeros(0)))
    y = A.spmv(torch.ones(5, dtype=torch.float64, device="cuda"), torch.full((7,), 9.0, dtype=torch.float64,
                                                                             device="cuda"))
    assert torch.equal(y.cpu(), torch.zeros(7, dtype=torch.float64))


@pytest.mark.parametrize("sigma", [8, 16, 32])
def test_fp32_product_error_bound(cuda, sigma):
    """FP32 values and vectors: no FP32 reference exists (the reference is
    FP64-only), so the error is measured against the FP64 oracle and must stay
    within the FP32 rounding bound (len_i + 2) * u32 * sum |a x| plus the
    input rounding of A and x."""
    import torch

    rp, ci, v = matgen.random_csr(4000, 4000, seed=sigma, max_len=40, empty_frac=0.05, long_rows=[(1, 3000)])
    x = np.random.default_rng(1).uniform(-1, 1, 4000)
    A = Csr5.from_csr((4000, 4000, rp, ci, v), sigma=sigma, keep_f32=True)
    y32 = A.spmv(to_dev(x, torch.float32)).double().cpu().numpy()
    ref = O.csr_spmv(rp, ci, v, x)
    u32 = np.finfo(np.float32).eps / 2
    lens = np.diff(rp.astype(np.int64))
    rows = np.repeat(np.arange(4000), lens)
    absum = np.bincount(rows, weights=np.abs(v * x[ci]), minlength=4000)
    assert (np.abs(y32 - ref) <= (lens + 4) * u32 * absum + 1e-30).all()
    rel = np.linalg.norm(y32 - ref) / np.linalg.norm(ref)
    assert rel < 1e-5


def test_fp32_requires_keep_f32(cuda):
    import torch

    rp, ci, v = matgen.random_csr(100, 100, 0)
    A = Csr5.from_csr((100, 100, rp, ci, v))
    with pytest.raises(ValueError, match="KEEP_F32"):
        A.spmv(torch.zeros(100, dtype=torch.float32, device="cuda"))


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 11, 16])
@pytest.mark.parametrize("sigma", [8, 16, 32])
def test_spmm_equals_columnwise_spmv(cuda, k, sigma):
    """One tile decode for k right-hand sides: integer inputs make it exact."""
    import torch

    rp, ci, v = matgen.random_csr(3000, 2500, seed=k, max_len=25, empty_frac=0.1, long_rows=[(4, 2400)],
                                  integer=True)
    X = np.random.default_rng(k).integers(-4, 5, (2500, k)).astype(np.float64)
    A = Csr5.from_csr((3000, 2500, rp, ci, v), sigma=sigma, keep_f32=True)
    Y = A.spmm(to_dev(X)).cpu().numpy()
    for j in range(k):
        np.testing.assert_array_equal(Y[:, j], O.csr_spmv(rp, ci, v, X[:, j].copy()))
    Y32 = A.spmm(to_dev(X, torch.float32)).double().cpu().numpy()
    np.testing.assert_array_equal(Y32, Y)  # small integers are exact in FP32 too


@pytest.mark.parametrize("rows", [1, 40, 300])
def test_spmm_sweep_with_fewer_chunks_than_xcds(cuda, rows):
    """Multi-RHS products (k > 2 take the ordered sweep) of matrices with one
    to a few chunks, so that most XCD ranges are empty: repeated products
    (and products of another matrix in between) must each be exact, which
    they are only if every launch leaves the chunk counters at zero."""
    rng = np.random.default_rng(rows)
    rp, ci, v = matgen.random_csr(rows, 50, seed=rows, max_len=6, integer=True)
    A = Csr5.from_csr((rows, 50, rp, ci, v), sigma=4)
    B = Csr5.from_csr((rows, 50, rp, ci, v), sigma=8)
    for rep in range(3):
        for M in (A, B):
            for k in (3, 8, 16):
                X = rng.integers(-4, 5, (50, k)).astype(np.float64)
                Y = M.spmm(to_dev(X)).cpu().numpy()
                for j in range(k):
                    np.testing.assert_array_equal(Y[:, j], O.csr_spmv(rp, ci, v, X[:, j].copy()), err_msg=f"{rep} {k}")


@pytest.mark.parametrize("k", [2, 3, 8, 13, 16])
@pytest.mark.parametrize("sigma", [4, 7, 8])
@pytest.mark.parametrize("layout", ["row", "col", "col_ld"])
def test_spmm_short_rows_both_layouts(cuda, k, sigma, layout):
    """Tiles of short rows (the 7-point operator on a 40^3 grid) take the
    per-side fast path for sigma 7 and 8 and a mix with the general path for
    sigma 4; integer inputs make FP64 and FP32 exact. Column-major blocks, also
    with a leading dimension beyond the column length, give column-major Y."""
    import torch

    side = 40
    rp, ci, v = matgen.poisson3d(side)
    n = side ** 3
    X = np.random.default_rng(k * 100 + sigma).integers(-4, 5, (n, k)).astype(np.float64)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma, keep_f32=True)
    want = np.stack([O.csr_spmv(rp, ci, v, X[:, j].copy()) for j in range(k)], 1)
    for dt in (torch.float64, torch.float32):
        Xd = to_dev(X, dt)
        if layout == "col":
            Xd = Xd.t().contiguous().t()
        elif layout == "col_ld":
            buf = torch.full((k, n + 7), float("nan"), dtype=dt, device="cuda")
            buf[:, :n] = Xd.t()
            Xd = buf[:, :n].t()
        Y = A.spmm(Xd)
        assert (Y.stride(0) == 1) == (layout != "row")
        np.testing.assert_array_equal(Y.double().cpu().numpy(), want)
    if layout == "col_ld":
        Ybuf = torch.full((k, n + 3), -1.0, dtype=torch.float64, device="cuda")
        A.spmm(to_dev(X).t().contiguous().t(), Ybuf[:, :n].t())
        np.testing.assert_array_equal(Ybuf[:, :n].t().cpu().numpy(), want)
        assert (Ybuf[:, n:] == -1.0).all()  # nothing written past the column length


@pytest.mark.parametrize("sigma", SIGMAS)
def test_transpose_product(cuda, sigma):
    rp, ci, v = matgen.random_csr(2000, 1700, seed=sigma, max_len=15, empty_frac=0.1, long_rows=[(2, 1600)],
                                  integer=True)
    x = np.random.default_rng(2).integers(-5, 6, 2000).astype(np.float64)
    A = Csr5.from_csr((2000, 1700, rp, ci, v), sigma=sigma)
    y = A.spmv_transpose(to_dev(x)).cpu().numpy()
    rows = np.repeat(np.arange(2000), np.diff(rp.astype(np.int64)))
    exp = np.bincount(ci, weights=v * x[rows], minlength=1700)
    np.testing.assert_array_equal(y, exp)


@pytest.mark.parametrize("sigma", SIGMAS)
def test_sampled_outer_product(cuda, sigma):
    rp, ci, v = matgen.random_csr(1500, 1200, seed=sigma, max_len=15, empty_frac=0.1, long_rows=[(2, 1100)])
    rng = np.random.default_rng(4)
    u = rng.standard_normal(1500)
    w = rng.standard_normal(1200)
    A = Csr5.from_csr((1500, 1200, rp, ci, v), sigma=sigma)
    out = A.sampled_outer(to_dev(u), to_dev(w), scale=-1.0).cpu().numpy()
    rows = np.repeat(np.arange(1500), np.diff(rp.astype(np.int64)))
    np.testing.assert_array_equal(out, -1.0 * u[rows] * w[ci])


@pytest.mark.parametrize("sigma", [8, 16, 32])
def test_row_block_shares(cuda, sigma):
    """spmv<partial> on tile ranges: each block equals the oracle's block
    product, and the blocks add up to the full product."""
    rp, ci, v = matgen.random_csr(6000, 6000, seed=sigma, max_len=30, empty_frac=0.1, long_rows=[(50, 5000)])
    x = np.random.default_rng(3).uniform(-1, 1, 6000)
    A = Csr5.from_csr((6000, 6000, rp, ci, v), sigma=sigma)
    T = A.tile_count
    cuts = sorted(set([0, T // 3, T // 2, T - 1, T]))
    acc = np.zeros(6000)
    tile = 64 * sigma
    xd = to_dev(x)
    for b, e in zip(cuts[:-1], cuts[1:]):
        first, yb = A.spmv_block(xd, b, e)
        yb = yb.cpu().numpy()
        ob = O.csr_block_spmv(rp, ci, v, b * tile, min(e * tile, len(v)), x, first, len(yb))
        np.testing.assert_allclose(yb, ob, rtol=1e-13, atol=1e-13)
        acc[first: first + len(yb)] += yb
    np.testing.assert_allclose(acc, O.csr_spmv(rp, ci, v, x), rtol=1e-12, atol=1e-12)


def test_autotuned_sigma(cuda):
    rp, ci, v = matgen.poisson2d(128)
    A = Csr5.from_csr((16384, 16384, rp, ci, v), sigma=0)
    # the tuner picks among the supported tile heights near 1-4x the mean row
    # length (5 here) plus 4, 8 and 16
    assert A.sigma in (4, 5, 6, 7, 8, 10, 12, 14, 16, 20)
    x = np.random.default_rng(0).uniform(-1, 1, 16384)
    assert_spmv_close(A.spmv(to_dev(x)).cpu().numpy(), O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


# ------------------------------------------------ full benchmark sizes ---

def test_poisson3d_128_full_size_closed_form(cuda):
    """The 128^3 7-point operator (n = 2,097,152): A*1 equals the count of
    missing neighbours, exactly; a random x matches the oracle."""
    import torch

    side = 128
    rp, ci, v = matgen.poisson3d(side)
    n = side ** 3
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=16)
    y = A.spmv(torch.ones(n, dtype=torch.float64, device="cuda")).cpu().numpy()
    i = np.arange(side)
    edge = (i == 0).astype(int) + (i == side - 1)
    missing = edge[:, None, None] + edge[None, :, None] + edge[None, None, :]
    np.testing.assert_array_equal(y, missing.ravel().astype(float))
    x = np.random.default_rng(11).uniform(-1, 1, n)
    assert_spmv_close(A.spmv(to_dev(x)).cpu().numpy(), O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


@pytest.mark.parametrize("sigma", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("empty_frac", [0.0, 0.05])
def test_short_rows_staged_writes_exact(cuda, sigma, empty_frac):
    """Rows of 1-3 elements put most tiles on the general path, whose row sums
    plain products of tile height <= 8 stage in LDS and write per tile in one
    coalesced pass (tiles with empty rows, here 5% of the rows in one case,
    keep their direct stores): with
    integer values every summation order is exact, so y must equal the
    oracle's bit for bit, in FP64 and FP32, including rows that cross tiles
    and chunks, the padded last tile and a sliced form of the same matrix."""
    import torch

    rows, cols = 50_000, 40_000
    rng = np.random.default_rng(sigma)
    lens = rng.integers(1, 4, rows)  # 1-3 elements: clean tiles, most on the general path
    lens[rng.random(rows) < empty_frac] = 0
    lens[17], lens[rows - 1] = 900, 700  # rows across several tiles and chunks
    rp = np.zeros(rows + 1, np.uint32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(cols, int(L), replace=False)) for L in lens]).astype(np.uint32)
    v = rng.integers(-8, 9, len(ci)).astype(np.float64)
    x = rng.integers(-4, 5, cols).astype(np.float64)
    ref = O.csr_spmv(rp, ci, v, x)
    A = Csr5.from_csr((rows, cols, rp, ci, v), sigma=sigma, keep_f32=True)
    np.testing.assert_array_equal(A.spmv(to_dev(x)).cpu().numpy(), ref)
    np.testing.assert_array_equal(A.spmv(to_dev(x, torch.float32)).double().cpu().numpy(), ref)
    S = Csr5.from_csr((rows, cols, rp, ci, v), sigma=sigma, xcd_slices=True)
    np.testing.assert_array_equal(S.spmv(to_dev(x)).cpu().numpy(), ref)


@pytest.mark.parametrize("sliced", [False, True])
def test_host_pointer_product_equals_device_product(cuda, sliced):
    """dim_csr5_spmv_host (numpy vectors in host memory, staged through the
    handle) gives the device product's y bit for bit, FP64 and FP32, on
    repeated calls, and rejects wrong sizes and dtypes like the device path."""
    import torch

    rp, ci, v = matgen.random_csr(5000, 4000, seed=3, max_len=12, empty_frac=0.1, long_rows=[(9, 3000)])
    A = Csr5.from_csr((5000, 4000, rp, ci, v), sigma=8, keep_f32=True, xcd_slices=sliced)
    rng = np.random.default_rng(4)
    for _ in range(2):
        x = rng.uniform(-1, 1, 4000)
        y_host = A.spmv(x)
        assert isinstance(y_host, np.ndarray)
        np.testing.assert_array_equal(y_host, A.spmv(to_dev(x)).cpu().numpy())
        x32 = x.astype(np.float32)
        np.testing.assert_array_equal(A.spmv(x32), A.spmv(to_dev(x32)).cpu().numpy())
    y = np.empty(5000)
    assert A.spmv(x, y) is y
    with pytest.raises(ValueError, match="column count"):
        A.spmv(np.zeros(3999))
    with pytest.raises(TypeError):
        A.spmv(np.zeros(4000, dtype=np.int64))
    with pytest.raises(TypeError):
        A.spmv(x, np.empty(5000, dtype=np.float32))
    assert torch.cuda.is_available()


def test_powerlaw_4m_full_size(cuda):
    """The irregular power-law matrix at n = 4,000,000 against the oracle."""
    n = 4_000_000
    rp, ci, v = matgen.powerlaw(n, seed=2024)
    x = np.random.default_rng(7).uniform(-1, 1, n)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=16)
    y = A.spmv(to_dev(x)).cpu().numpy()
    assert_spmv_close(y, O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


@pytest.mark.parametrize("form", ["sigma16", "autotuned"])
def test_powerlaw_4m_full_size_fp32(cuda, form):
    """The power-law matrix at n = 4,000,000 in FP32 (the benchmark's FP32
    leg: values and vectors in FP32) against the FP64 oracle, in the plain
    form and in the autotuned one (the column-sliced form when the tuner
    picks it): every row within the FP32 rounding bound (len_i + 4) * u32 *
    sum_j |a_ij x_j| (input rounding of A and x included), the normwise error
    stated."""
    import torch

    n = 4_000_000
    rp, ci, v = matgen.powerlaw(n, seed=2024)
    x = np.random.default_rng(7).uniform(-1, 1, n)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=16 if form == "sigma16" else 0, keep_f32=True)
    y32 = A.spmv(to_dev(x, torch.float32)).double().cpu().numpy()
    ref = O.csr_spmv(rp, ci, v, x)
    u32 = np.finfo(np.float32).eps / 2
    lens = np.diff(rp.astype(np.int64))
    absum = np.bincount(np.repeat(np.arange(n), lens), weights=np.abs(v * x[ci]), minlength=n)
    bad = np.abs(y32 - ref) > (lens + 4) * u32 * absum + 1e-30
    assert not bad.any(), np.nonzero(bad)[0][:5]
    rel = np.linalg.norm(y32 - ref) / np.linalg.norm(ref)
    print(f"{form}: sigma {A.sigma}, slices {A.info.xcd_slices}, FP32 normwise error vs FP64 oracle {rel:.2e}")
    assert rel < 1e-6


@pytest.mark.parametrize("sigma", [4, 5, 6, 7, 8])
def test_column_codes_equal_plain_columns(cuda, sigma):
    """Stencil operators keep column codes (their elements lie on 7 or 5
    diagonals): products read the codes and rebuild each column from its
    row, with the same summation order as the 32-bit columns, so results
    are bitwise equal to the plain format's, FP64 and FP32, whole and by
    tile blocks; with integer inputs they equal the oracle exactly."""
    import torch

    for rp, ci, v, n in [(*matgen.poisson3d(40), 40 ** 3), (*matgen.poisson2d(300), 300 ** 2)]:
        A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma, keep_f32=True)
        B = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma, keep_f32=True, plain_columns=True)
        assert A.info.column_codes == 1 and B.info.column_codes == 0
        xi = np.random.default_rng(sigma).integers(-5, 6, n).astype(np.float64)
        np.testing.assert_array_equal(A.spmv(to_dev(xi)).cpu().numpy(), O.csr_spmv(rp, ci, v, xi))
        x = np.random.default_rng(sigma + 1).uniform(-1, 1, n)
        for dt in (torch.float64, torch.float32):
            assert torch.equal(A.spmv(to_dev(x, dt)), B.spmv(to_dev(x, dt)))
        T = A.info.tile_count
        for tb, te in [(0, T // 3), (T // 3, T), (1, 2)]:
            fa, ya = A.spmv_block(to_dev(x), tb, te)
            fb, yb = B.spmv_block(to_dev(x), tb, te)
            assert fa == fb and torch.equal(ya, yb)


@pytest.mark.parametrize("sigma", [4, 5, 6, 7, 8])
def test_lookahead_across_chunk_lengths(cuda, sigma):
    """A row that crosses from a chunk's last tile into the next chunk's
    first tile is finished by the earlier chunk (look-ahead); the pipelined
    kernels load its elements in tile t1 during the chunk's last steps, into
    the stage slot that no prefetch needs any more. The grids here give
    chunks of one to five tiles (odd and even counts, so both slots of the
    two-deep pipeline), with column codes and with 32-bit columns; integer
    inputs must give the oracle's result exactly, FP64 and FP32."""
    import torch

    for n in (60, 73, 87, 101):
        rp, ci, _ = matgen.poisson3d(n)
        rows = n ** 3
        rng = np.random.default_rng(n * 10 + sigma)
        v = rng.integers(1, 5, len(ci)).astype(np.float64) * rng.choice([-1.0, 1.0], len(ci))
        xi = rng.integers(-5, 6, rows).astype(np.float64)
        expect = O.csr_spmv(rp, ci, v, xi)
        for plain in (False, True):
            A = Csr5.from_csr((rows, rows, rp, ci, v), sigma=sigma, keep_f32=True, plain_columns=plain)
            assert A.info.column_codes == (0 if plain else 1)
            np.testing.assert_array_equal(A.spmv(to_dev(xi)).cpu().numpy(), expect, err_msg=f"n={n} plain={plain}")
            y32 = A.spmv(to_dev(xi, torch.float32)).cpu().numpy()
            np.testing.assert_array_equal(y32.astype(np.float64), expect, err_msg=f"f32 n={n} plain={plain}")


@pytest.mark.parametrize("shape", ["wide", "tiny", "one_row"])
def test_column_codes_rectangular_and_small(cuda, shape):
    """Codes on rectangular banded matrices (diagonals far off the main one)
    and on matrices smaller than one tile equal the plain columns bit for
    bit."""
    import torch

    if shape == "wide":  # row i holds columns i, i + 1000, i + 2000 (3 diagonals)
        rows, cols, offs = 1000, 3000, [0, 1000, 2000]
    elif shape == "tiny":
        rows, cols, offs = 37, 40, [0, 3]
    else:
        rows, cols, offs = 1, 500, [0, 7, 499]
    rp = np.arange(rows + 1, dtype=np.uint32) * len(offs)
    ci = (np.arange(rows)[:, None] + np.asarray(offs)[None, :]).ravel().astype(np.uint32)
    v = np.random.default_rng(rows).standard_normal(len(ci))
    x = np.random.default_rng(cols).uniform(-1, 1, cols)
    for sigma in (4, 6, 7, 8):
        A = Csr5.from_csr((rows, cols, rp, ci, v), sigma=sigma, keep_f32=True)
        B = Csr5.from_csr((rows, cols, rp, ci, v), sigma=sigma, keep_f32=True, plain_columns=True)
        assert A.info.column_codes == 1
        ya = A.spmv(to_dev(x))
        assert torch.equal(ya, B.spmv(to_dev(x)))
        assert torch.equal(A.spmv(to_dev(x, torch.float32)), B.spmv(to_dev(x, torch.float32)))
        assert_spmv_close(ya.cpu().numpy(), O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


def test_column_codes_in_cg_and_when_absent(cuda):
    """CG's fused product reads the codes too (same iterates bit for bit as
    the plain format). Matrices with more than 16 diagonals or an empty row
    keep no codes."""
    import torch

    from diploma_research_amd import conjugate_gradient

    n = 120 * 120
    rp, ci, v = matgen.poisson2d(120)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=5)
    B = Csr5.from_csr((n, n, rp, ci, v), sigma=5, plain_columns=True)
    assert A.info.column_codes == 1
    b = to_dev(np.random.default_rng(3).uniform(-1, 1, n))
    xa, ra = conjugate_gradient(A, b, rtol=1e-10)
    xb, rb = conjugate_gradient(B, b, rtol=1e-10)
    assert ra.iterations == rb.iterations and torch.equal(xa, xb)
    rp2, ci2, v2 = matgen.random_csr(3000, 3000, seed=5, max_len=8)
    assert Csr5.from_csr((3000, 3000, rp2, ci2, v2), sigma=8).info.column_codes == 0
    lens = np.diff(rp.astype(np.int64))  # the same operator with row 7 emptied
    lens[7] = 0
    keep = np.ones(len(ci), bool)
    keep[rp[7]:rp[8]] = False
    rp3 = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    C = Csr5.from_csr((n, n, rp3, ci[keep], v[keep]), sigma=5)
    assert C.info.column_codes == 0


def _slice_case(name):
    if name == "powerlaw":
        n = 200_000
        rp, ci, v = matgen.powerlaw(n, seed=11)
        return n, n, rp, ci, v
    if name == "wide":
        rp, ci, v = matgen.random_csr(5000, 300_000, seed=12, max_len=40, empty_frac=0.1, long_rows=[(7, 250_000)])
        return 5000, 300_000, rp, ci, v
    if name == "empty_rows":
        rp, ci, v = matgen.random_csr(20_000, 20_000, seed=13, max_len=30, empty_frac=0.5, long_rows=[(3, 15_000)])
        return 20_000, 20_000, rp, ci, v
    if name == "few_cols":  # fewer columns than slices: some slices are empty
        rp, ci, v = matgen.random_csr(3000, 5, seed=14, max_len=5, empty_frac=0.2)
        return 3000, 5, rp, ci, v
    raise KeyError(name)


@pytest.mark.parametrize("name", ["powerlaw", "wide", "empty_rows", "few_cols"])
@pytest.mark.parametrize("sigma", [4, 8, 16])
def test_xcd_slices_match_oracle(cuda, name, sigma):
    """The column-sliced product (the elements of each column slice stacked
    as row blocks, partial products added per row in slice order) against the
    oracle: FP64 within the parity gate, FP32 within its rounding bound,
    repeat calls bitwise equal, spmv_timed equal to spmv."""
    import torch

    rows, cols, rp, ci, v = _slice_case(name)
    x = np.random.default_rng(sigma).uniform(-1, 1, cols)
    A = Csr5.from_csr((rows, cols, rp, ci, v), sigma=sigma, keep_f32=True, xcd_slices=True)
    assert A.info.xcd_slices == 8 and A.info.xcd_sigma == sigma
    ref = O.csr_spmv(rp, ci, v, x)
    y = A.spmv(to_dev(x))
    assert_spmv_close(y.cpu().numpy(), ref, rp, ci, v, x)
    assert torch.equal(A.spmv(to_dev(x)), y)
    yt = torch.full_like(y, float("nan"))
    A.spmv_timed(to_dev(x), yt)
    assert torch.equal(yt, y)
    y32 = A.spmv(to_dev(x, torch.float32)).double().cpu().numpy()
    u32 = np.finfo(np.float32).eps / 2
    lens = np.diff(rp.astype(np.int64))
    absum = np.bincount(np.repeat(np.arange(rows), lens), weights=np.abs(v * x[ci]), minlength=rows)
    assert (np.abs(y32 - ref) <= (lens + 4) * u32 * absum + 1e-30).all()


def test_xcd_slices_empty_matrix(cuda):
    import torch

    A = Csr5.from_csr((7, 5, np.zeros(8, np.uint32), np.zeros(0, np.uint32), np.zeros(0)), xcd_slices=True)
    y = A.spmv(torch.ones(5, dtype=torch.float64, device="cuda"), torch.full((7,), 9.0, dtype=torch.float64,
                                                                             device="cuda"))
    assert torch.equal(y.cpu(), torch.zeros(7, dtype=torch.float64))


def test_xcd_slices_integer_exact_and_set_values(cuda):
    """Integer inputs: the sliced sums are exact. Replacing the values drops
    the sliced form, and products then follow the new values."""
    import torch

    rp, ci, v = matgen.random_csr(6000, 90_000, seed=21, max_len=60, empty_frac=0.1, integer=True)
    x = np.random.default_rng(2).integers(-3, 4, 90_000).astype(np.float64)
    A = Csr5.from_csr((6000, 90_000, rp, ci, v), sigma=8, xcd_slices=True)
    np.testing.assert_array_equal(A.spmv(to_dev(x)).cpu().numpy(), O.csr_spmv(rp, ci, v, x))
    v2 = v * 2 - 1
    A.set_values(to_dev(v2))
    info = N.Csr5Info()
    N.check(N.lib().dim_csr5_get_info(A._h, ctypes.byref(info)))
    assert info.xcd_slices == 0
    np.testing.assert_array_equal(A.spmv(to_dev(x)).cpu().numpy(), O.csr_spmv(rp, ci, v2, x))


def test_xcd_slices_autotuned_on_powerlaw_4m(cuda):
    """sigma = 0 on the 4M power-law matrix (x = 32 MB, random columns): the
    autotuner weighs the sliced form against the unsliced one; the product it
    keeps matches the oracle either way."""
    n = 4_000_000
    rp, ci, v = matgen.powerlaw(n, seed=2024)
    x = np.random.default_rng(7).uniform(-1, 1, n)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=0)
    print("xcd_slices", A.info.xcd_slices, "xcd_sigma", A.info.xcd_sigma, "sigma", A.sigma)
    y = A.spmv(to_dev(x)).cpu().numpy()
    assert_spmv_close(y, O.csr_spmv(rp, ci, v, x), rp, ci, v, x)


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("sigma", [4, 7, 16, 32])
def test_gpu_conversion_equals_host_conversion(cuda, case, sigma):
    """CSR -> CSR5 on the GPU (from device CSR arrays) builds exactly the
    format of the host builder, so products are bitwise identical."""
    import torch

    c = CASES[case]
    rp, ci, v = matgen.random_csr(c["rows"], c["cols"], seed=200 + case, max_len=c["max_len"],
                                  empty_frac=c["empty_frac"], long_rows=c["long_rows"])
    H = Csr5.from_csr((c["rows"], c["cols"], rp, ci, v), sigma=sigma, keep_f32=True)
    D = Csr5.from_csr_device(c["rows"], c["cols"], to_dev(rp.astype(np.int32)), to_dev(ci.astype(np.int32)),
                             to_dev(v), sigma=sigma, keep_f32=True)
    fh, fd = H.format_arrays(), D.format_arrays()
    for key in fh:
        np.testing.assert_array_equal(fd[key], fh[key], err_msg=key)
    assert D.info.n_dirty_tiles == H.info.n_dirty_tiles and D.info.n_empty_rows == H.info.n_empty_rows
    x = to_dev(np.random.default_rng(case).uniform(-1, 1, c["cols"]))
    assert torch.equal(D.spmv(x), H.spmv(x))
    assert torch.equal(D.spmv(x.float()), H.spmv(x.float()))


def test_gpu_conversion_rejects_bad_csr(cuda):
    rp = np.array([0, 2, 1, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    with pytest.raises(ValueError, match="non-decreasing"):
        Csr5.from_csr_device(3, 3, to_dev(rp), to_dev(ci), to_dev(np.ones(3)))
    rp = np.array([0, 1, 2, 3], np.int32)
    ci = np.array([0, 1, 7], np.int32)
    with pytest.raises(ValueError, match="out of range"):
        Csr5.from_csr_device(3, 3, to_dev(rp), to_dev(ci), to_dev(np.ones(3)))
