"""GPU conjugate gradient against the oracle's textbook CG (same stop rule):
iteration counts within +-2, residual histories, reproducibility, error
paths, multi-RHS and the FP32-refined mode."""
import numpy as np
import pytest

from gpu_util import to_dev
from diploma_research_amd import Csr5, DimError, conjugate_gradient, conjugate_gradient_multi, matgen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _problem(rp, ci, v, n, seed):
    xs = np.random.default_rng(seed).uniform(-1, 1, n)
    b = O.csr_spmv(rp, ci, v, xs)
    return xs, b


@pytest.mark.parametrize("sigma", [5, 6, 8, 16, 32])
def test_cg_poisson2d_256_matches_oracle(cuda, sigma):
    """2-D 5-point Laplacian 256^2 (the CLI workload), rtol 1e-10."""
    rp, ci, v = matgen.poisson2d(256)
    n = 65536
    xs, b = _problem(rp, ci, v, n, 3)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma)
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-10, max_iters=5000, history=True)
    oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=5000, history=True)
    assert res.converged and oc["converged"]
    assert abs(res.iterations - oc["iterations"]) <= 2
    assert res.relres <= 1e-10 and res.relres_true <= 1e-9
    m = min(len(res.history), len(oc["history"]))
    # the residual histories agree while rounding differences stay small
    np.testing.assert_allclose(res.history[:40], oc["history"][:40], rtol=1e-8)
    assert np.abs(np.log10(res.history[:m]) - np.log10(oc["history"][:m])).max() < 0.5
    np.testing.assert_allclose(x.cpu().numpy(), xs, atol=1e-7)


@pytest.mark.parametrize("sigma", [16, 7])
def test_cg_poisson3d_128_full_size(cuda, sigma):
    """The benchmark operator: 128^3 7-point Poisson, b = A x*, rtol 1e-8,
    max 5000 iterations: same iteration count as the oracle within +-2
    (sigma 7: the column-coded, two-deep pipelined product with look-ahead)."""
    rp, ci, v = matgen.poisson3d(128)
    n = 128 ** 3
    xs, b = _problem(rp, ci, v, n, 42)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma)
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-8, max_iters=5000)
    oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-8, maxit=5000, nthreads=16)
    assert res.converged and oc["converged"]
    assert abs(res.iterations - oc["iterations"]) <= 2, (res.iterations, oc["iterations"])
    assert res.relres_true <= 2e-8
    err = np.linalg.norm(x.cpu().numpy() - oc["x"]) / np.linalg.norm(oc["x"])
    assert err < 1e-6


@pytest.mark.parametrize("sigma", [7, 16])
def test_cg_beyond_infinity_cache_matches_oracle(cuda, sigma):
    """168^3 Poisson (4.7M rows, 33M stored elements): the matrix and the
    solver's four vectors take 0.44 GB (sigma 7, column codes) or 0.56 GB
    (sigma 16), more than the 256 MB Infinity Cache, so the product fused
    with p.q streams the matrix with non-temporal loads. Same iteration
    count as the oracle's CG within +-2, and the same solution."""
    rp, ci, v = matgen.poisson3d(168)
    n = 168 ** 3
    xs, b = _problem(rp, ci, v, n, 168)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma)
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-8, max_iters=5000)
    oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-8, maxit=5000, nthreads=16)
    assert res.converged and oc["converged"]
    assert abs(res.iterations - oc["iterations"]) <= 2, (res.iterations, oc["iterations"])
    assert res.relres_true <= 2e-8
    err = np.linalg.norm(x.cpu().numpy() - oc["x"]) / np.linalg.norm(oc["x"])
    assert err < 1e-6


def _banded_blocks(n, block, width, seed):
    """Symmetric, diagonally dominant matrix whose rows alternate, block by
    block, between a band of half-width `width` and a tridiagonal band:
    2 * width + 1 diagonals in all (column codes), wide rows that leave lanes
    without a row start (the general path) next to short rows (the fast
    path)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(n):
        w = width if (i // block) % 2 == 0 else 1
        for d in range(1, w + 1):
            j = i + d
            if j < n:
                a = -rng.uniform(0.1, 1.0)
                rows += [i, j]
                cols += [j, i]
                vals += [a, a]
    rows = np.array(rows, dtype=np.int64)
    cols = np.array(cols, dtype=np.int64)
    vals = np.array(vals)
    diag = 1.0 + np.bincount(rows, weights=np.abs(vals), minlength=n)
    rows = np.concatenate([rows, np.arange(n)])
    cols = np.concatenate([cols, np.arange(n)])
    vals = np.concatenate([vals, diag])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, dtype=np.uint32)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return rp, cols.astype(np.uint32), vals


@pytest.mark.parametrize("sigma", [5, 6])
def test_cg_coded_wide_and_short_rows(cuda, sigma):
    """Column-coded CG product with the fused dot where tiles of wide rows
    (general path) alternate with tiles of short rows (fast path, x at the
    rows taken from the diagonal elements' gathers): a row carried out of a
    general-path tile, or with its diagonal in an earlier tile, falls back to
    loading x at the rows. Iteration count within +-2 of the oracle's CG and
    the same solution."""
    n = 24000
    rp, ci, v = _banded_blocks(n, 700, 7, sigma)
    xs, b = _problem(rp, ci, v, n, 11)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=sigma)
    assert A.info.column_codes == 1
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-10, max_iters=5000)
    oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=5000)
    assert res.converged and oc["converged"]
    assert abs(res.iterations - oc["iterations"]) <= 2, (res.iterations, oc["iterations"])
    assert res.relres_true <= 1.01e-10
    np.testing.assert_allclose(x.cpu().numpy(), oc["x"], rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("n", [73, 87])
def test_cg_coded_products_of_short_chunks(cuda, n):
    """CG's product with the fused p.q on grids whose chunks hold two and
    three tiles (look-ahead elements loaded into either pipeline slot):
    iteration count within +-2 of the oracle's CG."""
    rp, ci, v = matgen.poisson3d(n)
    rows = n ** 3
    xs, b = _problem(rp, ci, v, rows, n)
    A = Csr5.from_csr((rows, rows, rp, ci, v), sigma=7)
    assert A.info.column_codes == 1
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-8, max_iters=5000)
    oc = O.Csr5(rows, rows, rp, ci, v).cg(b, rtol=1e-8, maxit=5000, nthreads=16)
    assert res.converged and oc["converged"]
    assert abs(res.iterations - oc["iterations"]) <= 2, (res.iterations, oc["iterations"])
    assert res.relres_true <= 2e-8


def test_cg_host_pointers_equal_device_solve(cuda):
    """dim_cg_host (b and x in host memory, x0 nonzero, with history) returns
    the device solve's iterate, result and history bit for bit; a solve that
    stops at max_iters still hands back its iterate."""
    n = 48
    rp, ci, v = matgen.poisson2d(n)
    N = n * n
    A = Csr5.from_csr((N, N, rp, ci, v), sigma=0)
    rng = np.random.default_rng(21)
    b = rng.uniform(-1, 1, N)
    x0 = rng.uniform(-0.1, 0.1, N)
    xd, rd = conjugate_gradient(A, to_dev(b), to_dev(x0), rtol=1e-10, history=True)
    xh, rh = conjugate_gradient(A, b, x0.copy(), rtol=1e-10, history=True)
    assert isinstance(xh, np.ndarray)
    np.testing.assert_array_equal(xh, xd.cpu().numpy())
    assert (rh.iterations, rh.converged, rh.relres, rh.relres_true) == \
        (rd.iterations, rd.converged, rd.relres, rd.relres_true)
    np.testing.assert_array_equal(rh.history, rd.history)
    xc, rc = conjugate_gradient(A, b, None, rtol=1e-14, max_iters=7)
    xcd, _ = conjugate_gradient(A, to_dev(b), None, rtol=1e-14, max_iters=7)
    assert not rc.converged and rc.iterations == 7
    np.testing.assert_array_equal(xc, xcd.cpu().numpy())


def test_cg_repeated_solves_on_one_handle_equal_fresh_handles(cuda):
    """A handle keeps its last solve's work buffers and captured iterations
    for the next solve of the same shape. Solves in a row on one handle (the
    same x buffer again, another x buffer, another b, other right-hand-side
    counts, precisions and history lengths, then back) equal the same solves
    on fresh handles bit for bit."""
    import torch

    rp, ci, v = matgen.varcoef((40, 40), 0.5, seed=7)
    n = 1600
    csr = (n, n, rp, ci, v)
    A = Csr5.from_csr(csr, sigma=5, keep_f32=True)
    rng = np.random.default_rng(4)
    b1, b2 = to_dev(rng.uniform(-1, 1, n)), to_dev(rng.uniform(-1, 1, n))
    B3 = to_dev(rng.uniform(-1, 1, (n, 3)))
    xbuf = torch.zeros(n, dtype=torch.float64, device="cuda")

    def single(M, b, x=None, **kw):
        xs, r = conjugate_gradient(M, b, None if x is None else x.zero_(), rtol=1e-10, max_iters=3000, **kw)
        return xs.clone(), r

    steps = [
        lambda M: single(M, b1),
        lambda M: single(M, b2, xbuf),
        lambda M: single(M, b1, xbuf),                       # same x buffer: the captured iterations again
        lambda M: single(M, b1, history=True),
        lambda M: single(M, b2, precision="f32-refined"),
        lambda M: conjugate_gradient_multi(M, B3, rtol=1e-10, max_iters=3000),
        lambda M: single(M, b2, xbuf),
    ]
    for step in steps:
        x_kept, r_kept = step(A)
        x_new, r_new = step(Csr5.from_csr(csr, sigma=5, keep_f32=True))
        assert torch.equal(x_kept, x_new)
        rk = r_kept if isinstance(r_kept, list) else [r_kept]
        rn = r_new if isinstance(r_new, list) else [r_new]
        assert [r.iterations for r in rk] == [r.iterations for r in rn]
        assert all(r.converged for r in rk)


def test_cg_is_bitwise_reproducible(cuda):
    rp, ci, v = matgen.poisson3d(40)
    n = 40 ** 3
    _, b = _problem(rp, ci, v, n, 1)
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=8)
    bd = to_dev(b)
    x1, r1 = conjugate_gradient(A, bd, rtol=1e-10)
    x2, r2 = conjugate_gradient(A, bd, rtol=1e-10)
    assert r1.iterations == r2.iterations
    assert bool((x1 == x2).all())


def test_cg_diagonal_distinct_eigenvalues(cuda):
    n = 4096
    d = np.repeat([1.0, 3.0, 7.0], [1365, 1365, 1366])
    A = Csr5.from_csr((n, n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d))
    b = np.random.default_rng(0).uniform(-1, 1, n)
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-12, max_iters=50)
    assert res.converged and res.iterations == 3
    np.testing.assert_allclose(x.cpu().numpy(), b / d, rtol=1e-12)


def test_cg_initial_guess_and_zero_rhs(cuda):
    import torch

    rp, ci, v = matgen.poisson2d(64)
    n = 4096
    xs, b = _problem(rp, ci, v, n, 5)
    A = Csr5.from_csr((n, n, rp, ci, v))
    x0 = to_dev(xs)  # exact solution as x0: converged before the first iteration
    x, res = conjugate_gradient(A, to_dev(b), x=x0.clone(), rtol=1e-8)
    assert res.converged and res.iterations == 0
    x, res = conjugate_gradient(A, torch.zeros(n, dtype=torch.float64, device="cuda"), rtol=1e-8)
    assert res.converged and res.iterations == 0 and float(x.abs().max()) == 0.0
    # a nonzero x0 continues from there, like the oracle
    x0 = np.random.default_rng(9).uniform(-1, 1, n)
    x, res = conjugate_gradient(A, to_dev(b), x=to_dev(x0), rtol=1e-10)
    oc = O.Csr5(n, n, rp, ci, v).cg(b, x0=x0, rtol=1e-10)
    assert res.converged and abs(res.iterations - oc["iterations"]) <= 2


def test_cg_breakdown_on_indefinite_matrix(cuda):
    n = 1000
    d = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    A = Csr5.from_csr((n, n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d))
    b = np.ones(n)
    with pytest.raises(DimError, match="breakdown"):
        conjugate_gradient(A, to_dev(b), rtol=1e-10)


def test_cg_iteration_cap(cuda):
    rp, ci, v = matgen.poisson2d(128)
    n = 128 * 128
    _, b = _problem(rp, ci, v, n, 2)
    A = Csr5.from_csr((n, n, rp, ci, v))
    x, res = conjugate_gradient(A, to_dev(b), rtol=1e-14, max_iters=37)
    assert not res.converged and res.iterations == 37
    oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-14, maxit=37)
    np.testing.assert_allclose(x.cpu().numpy(), oc["x"], rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("k", [2, 3, 8])
def test_cg_multi_rhs_equals_single(cuda, k):
    """k right-hand sides through one SpMM per iteration give, column by
    column, the iterates of k single solves: the same iteration count within
    two and the same solution to the solver tolerance (the coalesced multi-
    column reductions add the partial sums in another order)."""
    import torch

    rp, ci, v = matgen.varcoef((24, 24, 24), 1.0, seed=3)
    n = 24 ** 3
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=16)
    B = np.random.default_rng(k).uniform(-1, 1, (n, k))
    X, res = conjugate_gradient_multi(A, to_dev(B), rtol=1e-10, max_iters=3000)
    for j in range(k):
        xj, rj = conjugate_gradient(A, to_dev(B[:, j].copy()), rtol=1e-10, max_iters=3000)
        assert res[j].converged and abs(res[j].iterations - rj.iterations) <= 2
        assert res[j].relres_true <= 1.01e-10
        np.testing.assert_allclose(X[:, j].cpu().numpy(), xj.cpu().numpy(), rtol=1e-7, atol=1e-9)


def test_cg_multi_rhs_more_blocks_than_resident(cuda):
    """96^3 Poisson with 8 right-hand sides: the vector passes run 1728
    blocks of 1024 threads, more than the GPU holds at once, so blocks of
    the last x/p pass start after its block (0, 0) has set the stop flag.
    Every column must still take its last x update: true residuals at the
    tolerance, and each column equal to its single-RHS solve."""
    rp, ci, v = matgen.poisson3d(96)
    n = 96 ** 3
    k = 8
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=7)
    B = np.random.default_rng(96).uniform(-1, 1, (n, k))
    X, res = conjugate_gradient_multi(A, to_dev(B), rtol=1e-10, max_iters=3000)
    for j in range(k):
        xj, rj = conjugate_gradient(A, to_dev(B[:, j].copy()), rtol=1e-10, max_iters=3000)
        assert res[j].converged and abs(res[j].iterations - rj.iterations) <= 2
        assert res[j].relres_true <= 1.01e-10, (j, res[j].relres_true)
        np.testing.assert_allclose(X[:, j].cpu().numpy(), xj.cpu().numpy(), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("k", [1, 4])
def test_cg_fp32_refined_reaches_fp64_accuracy(cuda, k):
    """FP32 iteration with FP64 residual replacement reaches relative
    residual 1e-10 on a variable-coefficient operator (lognormal(0,1)
    diffusivity); the error against the FP64 oracle solution is stated."""
    rp, ci, v = matgen.varcoef((32, 32, 32), 1.0, seed=11)
    n = 32 ** 3
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=16, keep_f32=True)
    B = np.random.default_rng(7).uniform(-1, 1, (n, k))
    X, res = conjugate_gradient_multi(A, to_dev(B), rtol=1e-10, max_iters=20000, precision="f32-refined")
    O64 = O.Csr5(n, n, rp, ci, v)
    for j in range(k):
        oc = O64.cg(B[:, j].copy(), rtol=1e-10, maxit=20000)
        assert res[j].converged, res[j]
        assert res[j].relres_true <= 1.01e-10
        assert res[j].refinements >= 1
        # FP32 iterations cost some extra steps but stay close to FP64 CG
        assert res[j].iterations <= 1.5 * oc["iterations"] + 20, (res[j].iterations, oc["iterations"])
        err = np.linalg.norm(X[:, j].cpu().numpy() - oc["x"]) / np.linalg.norm(oc["x"])
        assert err < 1e-6, err


def test_cg_fp32_refined_k8_varcoef_lognormal_64(cuda):
    """The mixed-precision benchmark's shape at 64^3: FP32-refined CG with
    k = 8 right-hand sides (the FP32 tile kernel for 8 sides, the 8-column
    vector kernels and FP64 residual replacement) on the lognormal(0,1)
    variable-coefficient operator, with the autotuned format the benchmark
    uses. Every column against the oracle's FP64 CG: converged, FP64 true
    relative residual <= 1.01e-10, iterations <= 1.5x the oracle's + 20, and
    the solution's error against the oracle's x stated and bounded."""
    side, k = 64, 8
    rp, ci, v = matgen.varcoef((side, side, side), 1.0, seed=4)
    n = side ** 3
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=0, keep_f32=True)
    B = np.random.default_rng(5).uniform(-1, 1, (n, k))
    X, res = conjugate_gradient_multi(A, to_dev(B), rtol=1e-10, max_iters=20000, precision="f32-refined")
    Xh = X.cpu().numpy()
    O64 = O.Csr5(n, n, rp, ci, v)
    errs = []
    for j in range(k):
        oc = O64.cg(B[:, j].copy(), rtol=1e-10, maxit=20000, nthreads=16)
        assert oc["converged"]
        assert res[j].converged, (j, res[j])
        assert res[j].relres_true <= 1.01e-10, (j, res[j].relres_true)
        assert res[j].iterations <= 1.5 * oc["iterations"] + 20, (j, res[j].iterations, oc["iterations"])
        errs.append(np.linalg.norm(Xh[:, j] - oc["x"]) / np.linalg.norm(oc["x"]))
    print(f"sigma {A.sigma}, codes {A.info.column_codes}, refinements {res[0].refinements}, "
          f"iterations {[r.iterations for r in res]}, rel err vs oracle x {max(errs):.2e}")
    assert max(errs) < 1e-6, errs


def test_cg_fp32_refined_resumes_with_restarted_direction(cuda):
    """Residual replacement only when the FP32 recursion says converged (a
    negligible refine_delta), on a high-contrast operator (lognormal(0,2)):
    the FP64 residual is then still above the tolerance, so columns resume,
    each from p = r (the direction a converged column stopped updating is
    stale). Every column converges to the FP64 tolerance within twice the
    oracle's iterations + 50, after more than one replacement."""
    side, k = 40, 4
    rp, ci, v = matgen.varcoef((side, side, side), 2.0, seed=21)
    n = side ** 3
    A = Csr5.from_csr((n, n, rp, ci, v), sigma=7, keep_f32=True)
    B = np.random.default_rng(22).uniform(-1, 1, (n, k))
    X, res = conjugate_gradient_multi(A, to_dev(B), rtol=1e-10, max_iters=40000, precision="f32-refined",
                                      refine_delta=1e-30)
    O64 = O.Csr5(n, n, rp, ci, v)
    for j in range(k):
        oc = O64.cg(B[:, j].copy(), rtol=1e-10, maxit=40000, nthreads=16)
        assert res[j].converged and res[j].relres_true <= 1.01e-10, (j, res[j])
        assert res[j].iterations <= 2 * oc["iterations"] + 50, (j, res[j].iterations, oc["iterations"])
    assert res[0].refinements >= 2, res[0]
