"""Batched CG (many uneven SPD systems in one launch sequence) and the
hand-written adjoint behind torch.autograd: checked against the oracle's CG
per system, dense FP64 solves, and central finite differences."""
import numpy as np
import pytest

from gpu_util import assert_spmv_close, to_dev
from diploma_research_amd import Csr5, conjugate_gradient, matgen
from diploma_research_amd.csr5 import conjugate_gradient_batched
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _systems(sides, sigma_log, seed):
    out = []
    for i, s in enumerate(sides):
        rp, ci, v = matgen.varcoef((s, s), sigma_log, seed=seed + i)
        out.append((s * s, rp, ci, v))
    return out


def _iters_close(a, b):
    """CG iteration counts of two FP64 solves that reduce their dot products
    in different orders: equal to within 2, or 0.25% of a long solve (over
    thousands of iterations the rounding moves the stop by a few steps)."""
    return abs(a - b) <= max(2, b // 400)


def _dense(rp, ci, v, n):
    D = np.zeros((n, n))
    D[np.repeat(np.arange(n), np.diff(rp.astype(np.int64))), ci] = v
    return D


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("sigma", [0, 5, 16])
def test_batched_cg_matches_oracle_per_system(cuda, sigma, persistent):
    sides = [11, 37, 5, 60, 23, 48]
    systems = _systems(sides, 0.5, seed=1)
    A = Csr5.from_csr_batched(systems, sigma=sigma)
    assert A.n_systems == len(sides)
    rng = np.random.default_rng(0)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    X, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=5000,
                                        persistent=persistent)
    X = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=5000)
        assert r.converged and _iters_close(r.iterations, oc["iterations"]), (r.iterations, oc["iterations"])
        assert r.relres_true <= 2e-10
        np.testing.assert_allclose(X[off: off + n], oc["x"], rtol=1e-7, atol=1e-8)
        off += n


def test_batched_cg_more_blocks_than_resident(cuda):
    """Three large systems (2.3M rows, 563 blocks of 1024 threads, more than
    the GPU holds at once): blocks of the last x/p pass start after its block
    0 has set the stop flag, and must still take their last x update. Each
    system matches its own single-system solve."""
    systems = _systems([900, 700, 1000], 0.5, seed=9)
    A = Csr5.from_csr_batched(systems, sigma=0)
    rng = np.random.default_rng(3)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    X, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=20000)
    X = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        M = Csr5.from_csr((n, n, rp, ci, v), sigma=A.sigma)
        xs, rs = conjugate_gradient(M, to_dev(b), rtol=1e-10, max_iters=20000)
        assert r.converged and _iters_close(r.iterations, rs.iterations), (r.iterations, rs.iterations)
        assert r.relres_true <= 1.01e-10, r.relres_true
        np.testing.assert_allclose(X[off: off + n], xs.cpu().numpy(), rtol=1e-7, atol=1e-8)
        off += n


def test_batched_cg_beyond_infinity_cache_matches_single_solves(cuda):
    """Four systems of 8.0M rows in all (40M stored elements): the matrix and
    the solver's four vectors take about 0.6 GB, more than the 256 MB
    Infinity Cache, so the batch's products stream the matrix with
    non-temporal loads (the single solves of each system fit and use
    ordinary loads). Each system matches its own single-system solve."""
    systems = _systems([1500, 1300, 1700, 1100], 0.5, seed=21)
    A = Csr5.from_csr_batched(systems, sigma=5)
    rng = np.random.default_rng(8)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    X, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=30000)
    X = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        M = Csr5.from_csr((n, n, rp, ci, v), sigma=5)
        xs, rs = conjugate_gradient(M, to_dev(b), rtol=1e-10, max_iters=30000)
        assert r.converged and _iters_close(r.iterations, rs.iterations), (r.iterations, rs.iterations)
        assert r.relres_true <= 1.01e-10, r.relres_true
        np.testing.assert_allclose(X[off: off + n], xs.cpu().numpy(), rtol=1e-7, atol=1e-8)
        off += n


def test_batched_repeated_solves_on_one_handle_equal_fresh_handles(cuda):
    """A batch handle keeps its block tables, buffers and captured iterations
    from its first solve. Solves in a row on one handle (another b, the same
    x buffer again, another x buffer, the transposed product, the persistent
    kernel, other batch lengths) equal the same solves on fresh handles bit
    for bit; the transposed product adds with atomics in no fixed order, so
    that solve matches to the solver's tolerance."""
    import torch

    systems = _systems([21, 9, 33, 14], 0.5, seed=70)
    rng = np.random.default_rng(70)
    N = sum(n for n, *_ in systems)
    b1, b2 = to_dev(rng.uniform(-1, 1, N)), to_dev(rng.uniform(-1, 1, N))
    xbuf = torch.zeros(N, dtype=torch.float64, device="cuda")
    A = Csr5.from_csr_batched(systems, sigma=5)

    def run(M, b, x=None, **kw):
        X, res = conjugate_gradient_batched(M, b, None if x is None else x.zero_(), rtol=1e-11, max_iters=4000,
                                            **kw)
        return X.clone(), res

    steps = [
        (True, lambda M: run(M, b1)),
        (True, lambda M: run(M, b2, xbuf)),
        (True, lambda M: run(M, b1, xbuf)),          # same x buffer: the captured iterations again
        (False, lambda M: run(M, b2, transpose=True)),
        (True, lambda M: run(M, b1, persistent=True)),
        (True, lambda M: run(M, b2, xbuf, check_every=5)),
        (True, lambda M: run(M, b1, xbuf)),
    ]
    for exact, step in steps:
        x_kept, r_kept = step(A)
        x_new, r_new = step(Csr5.from_csr_batched(systems, sigma=5))
        assert all(r.converged for r in r_kept)
        if exact:
            assert torch.equal(x_kept, x_new)
            assert [r.iterations for r in r_kept] == [r.iterations for r in r_new]
        else:
            np.testing.assert_allclose(x_kept.cpu().numpy(), x_new.cpu().numpy(), rtol=1e-8, atol=1e-10)
            assert all(_iters_close(a.iterations, b.iterations) for a, b in zip(r_kept, r_new))


def test_batched_products_and_values(cuda):
    import torch

    systems = _systems([9, 31, 17], 0.5, seed=4)
    A = Csr5.from_csr_batched(systems, sigma=8)
    x = np.random.default_rng(1).uniform(-1, 1, sum(n for n, *_ in systems))
    y = A.spmv(to_dev(x)).cpu().numpy()
    yt = A.spmv_transpose(to_dev(x)).cpu().numpy()
    off = 0
    for n, rp, ci, v in systems:
        ref = O.csr_spmv(rp, ci, v, x[off: off + n].copy())
        np.testing.assert_allclose(y[off: off + n], ref, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(yt[off: off + n], ref, rtol=1e-12, atol=1e-13)  # symmetric systems
        off += n
    # replace the values: 2*A
    vals = torch.from_numpy(np.concatenate([v for *_, v in systems]) * 2.0).cuda()
    A.set_values(vals)
    np.testing.assert_allclose(A.spmv(to_dev(x)).cpu().numpy(), 2 * y, rtol=1e-14, atol=1e-14)


def test_transposed_solve_matches_forward_on_spd(cuda):
    systems = _systems([25, 14], 0.5, seed=9)
    A = Csr5.from_csr_batched(systems, sigma=8)
    b = to_dev(np.random.default_rng(3).uniform(-1, 1, 25 * 25 + 14 * 14))
    X1, r1 = conjugate_gradient_batched(A, b, rtol=1e-12)
    X2, r2 = conjugate_gradient_batched(A, b, rtol=1e-12, transpose=True)
    assert all(r.converged for r in r1 + r2)
    np.testing.assert_allclose(X2.cpu().numpy(), X1.cpu().numpy(), rtol=1e-9, atol=1e-11)


def test_batched_breakdown_is_reported(cuda):
    from diploma_research_amd import DimError

    n = 50
    d = np.where(np.arange(n) % 3 == 0, -1.0, 2.0)
    good = _systems([6], 0.5, seed=2)[0]
    bad = (n, np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d)
    A = Csr5.from_csr_batched([good, bad], sigma=8)
    with pytest.raises(DimError, match="breakdown"):
        conjugate_gradient_batched(A, to_dev(np.ones(36 + n)), rtol=1e-10)


# ------------------------------------------------------------- adjoint ---

def _loss_dense(systems, vals_list, bs):
    L = 0.0
    for (n, rp, ci, _), v, b in zip(systems, vals_list, bs):
        x = np.linalg.solve(_dense(rp, ci, v, n), b)
        L += 0.5 * float(x @ x)
    return L


@pytest.mark.parametrize("symmetric", [True, False])
def test_adjoint_gradient_matches_dense_reference_and_finite_differences(cuda, symmetric):
    """L = 0.5 ||x||^2, x = A^-1 b over a batch of uneven 2-D variable-
    coefficient systems. dL/db and dL/dvals from the adjoint (CG with A's own
    products, or with the transposed kernel when symmetric=False, + sampled
    outer product) against (1) the dense FP64 reference gradient and
    (2) fourth-order central finite differences, h = 1e-3, on 64 random
    entries, perturbing a_ij and a_ji together (the systems are symmetric, so
    the gradient of that parametrisation is g_ij + g_ji)."""
    import torch

    from diploma_research_amd.autograd import solve

    sides = [7, 12, 9, 15, 10, 13]
    systems = _systems(sides, 0.5, seed=21)
    A = Csr5.from_csr_batched(systems, sigma=8)
    rng = np.random.default_rng(5)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    vals_np = np.concatenate([v for *_, v in systems])
    vals = torch.from_numpy(vals_np.copy()).cuda().requires_grad_(True)
    b = torch.from_numpy(np.concatenate(bs)).cuda().requires_grad_(True)
    x = solve(A, vals, b, rtol=1e-14, max_iters=20000, symmetric=symmetric)
    L = 0.5 * (x * x).sum()
    L.backward()
    gv = vals.grad.cpu().numpy()
    gb = b.grad.cpu().numpy()

    # (1) dense FP64 reference gradient
    off_v = off_r = 0
    ref_gv = np.zeros_like(gv)
    ref_gb = np.zeros_like(gb)
    for (n, rp, ci, v), bb in zip(systems, bs):
        D = _dense(rp, ci, v, n)
        xs = np.linalg.solve(D, bb)
        lam = np.linalg.solve(D.T, xs)
        rows = np.repeat(np.arange(n), np.diff(rp.astype(np.int64)))
        ref_gv[off_v: off_v + len(v)] = -lam[rows] * xs[ci]
        ref_gb[off_r: off_r + n] = lam
        off_v += len(v)
        off_r += n
    np.testing.assert_allclose(gb, ref_gb, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(gv, ref_gv, rtol=1e-9, atol=1e-12)

    # (2) fourth-order central finite differences, h = 1e-3, on 64 random
    # stored entries
    h = 1e-3
    vals_list = [v.copy() for *_, v in systems]
    starts = np.concatenate([[0], np.cumsum([len(v) for v in vals_list])])
    picks = rng.choice(len(vals_np), size=64, replace=False)
    worst = 0.0
    for e in picks:
        s = int(np.searchsorted(starts, e, side="right") - 1)
        n, rp, ci, v = systems[s]
        k = e - starts[s]
        i = int(np.searchsorted(rp.astype(np.int64), k, side="right") - 1)
        j = int(ci[k])
        mirror = [k]
        if i != j:
            row_j = slice(int(rp[j]), int(rp[j + 1]))
            kk = int(rp[j]) + int(np.nonzero(ci[row_j] == i)[0][0])
            mirror.append(kk)

        def x_at(d):
            vs = vals_list[s].copy()
            for m in mirror:
                vs[m] += d
            return np.linalg.solve(_dense(rp, ci, vs, n), bs[s])

        # only system s depends on the entry; L(a) - L(b) is evaluated as
        # 0.5 (x_a - x_b).(x_a + x_b) so that the difference is not lost to the
        # rounding of L itself
        def dl(xa, xb):
            return 0.5 * float((xa - xb) @ (xa + xb))

        fd = (8 * dl(x_at(h), x_at(-h)) - dl(x_at(2 * h), x_at(-2 * h))) / (12 * h)
        an = sum(gv[starts[s] + m] for m in mirror)
        rel = abs(fd - an) / max(abs(an), 1e-12)
        worst = max(worst, rel)
    assert worst <= 1e-6, worst


def test_adjoint_gradient_second_order_fd_h1e6_through_gpu_solves(cuda):
    """The adjoint's dL/dvals against second-order central differences with
    h = 1e-6 on 64 random stored entries, every evaluation a GPU CG solve
    (rtol 1e-15). [L(+h) - L(-h)] / 2h is evaluated exactly as
    -0.5 d.(x(+h) + x(-h)) with A(+h) d = E x(-h) (E the unit perturbation of
    a_ij and a_ji), the resolvent identity x(+h) - x(-h) = -2h A(+h)^-1 E
    x(-h): the solves' rounding is then not divided by h, and what is left
    is the O(h^2) truncation. Bound: 1e-6 relative on every entry."""
    import torch

    from diploma_research_amd.autograd import solve

    systems = _systems([20, 35, 28, 41], 0.5, seed=41)
    A = Csr5.from_csr_batched(systems, sigma=0)
    rng = np.random.default_rng(19)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    vals_np = np.concatenate([v for *_, v in systems])
    vals = torch.from_numpy(vals_np.copy()).cuda().requires_grad_(True)
    b = torch.from_numpy(np.concatenate(bs)).cuda()
    x = solve(A, vals, b, rtol=1e-14, max_iters=20000)
    (0.5 * (x * x).sum()).backward()
    gv = vals.grad.cpu().numpy()

    h = 1e-6
    starts = np.concatenate([[0], np.cumsum([len(v) for *_, v in systems])])
    singles = {}
    worst = 0.0
    for e in rng.choice(len(vals_np), size=64, replace=False):
        s = int(np.searchsorted(starts, e, side="right") - 1)
        n, rp, ci, v = systems[s]
        if s not in singles:
            singles[s] = Csr5.from_csr((n, n, rp, ci, v), sigma=A.sigma)
        M = singles[s]
        k = int(e - starts[s])
        i = int(np.searchsorted(rp.astype(np.int64), k, side="right") - 1)
        j = int(ci[k])
        mirror = [k]
        if i != j:
            row_j = slice(int(rp[j]), int(rp[j + 1]))
            mirror.append(int(rp[j]) + int(np.nonzero(ci[row_j] == i)[0][0]))
        bd = to_dev(bs[s])
        xs = {}
        for step in (-h, h):
            vv = v.copy()
            for m in mirror:
                vv[m] += step
            M.set_values(torch.from_numpy(vv).cuda())
            xs[step], r = conjugate_gradient(M, bd, rtol=1e-15, max_iters=20000)
            assert r.converged
        w = torch.zeros_like(bd)
        if i != j:
            w[i], w[j] = xs[-h][j], xs[-h][i]
        else:
            w[i] = xs[-h][i]
        d, r = conjugate_gradient(M, w, rtol=1e-15, max_iters=20000)  # M holds A(+h)
        assert r.converged
        fd = -0.5 * float(torch.dot(d, xs[h] + xs[-h]))
        an = sum(gv[starts[s] + m] for m in mirror)
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-300))
    assert worst <= 1e-6, worst


def test_adjoint_gradient_wrt_b_by_finite_differences_on_gpu(cuda):
    """FD through the GPU solver itself (every evaluation a batched CG)."""
    import torch

    from diploma_research_amd.autograd import solve

    systems = _systems([8, 11], 0.5, seed=31)
    A = Csr5.from_csr_batched(systems, sigma=8)
    vals = torch.from_numpy(np.concatenate([v for *_, v in systems])).cuda()
    n = 64 + 121
    b0 = np.random.default_rng(8).uniform(-1, 1, n)
    b = torch.from_numpy(b0.copy()).cuda().requires_grad_(True)
    x = solve(A, vals, b, rtol=1e-15, max_iters=20000)
    (0.5 * (x * x).sum()).backward()
    g = b.grad.cpu().numpy()
    h = 1e-6
    for idx in np.random.default_rng(2).choice(n, size=8, replace=False):
        xs_pm = []
        for sgn in (+1, -1):
            bb = b0.copy()
            bb[idx] += sgn * h
            xx, _ = conjugate_gradient_batched(A, to_dev(bb), rtol=1e-15, max_iters=20000)
            xs_pm.append(xx.cpu().numpy())
        fd = 0.5 * float((xs_pm[0] - xs_pm[1]) @ (xs_pm[0] + xs_pm[1])) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(abs(g[idx]), 1e-3), (idx, fd, g[idx])


def test_autograd_raises_when_a_solve_stops_at_max_iters(cuda):
    """The autograd path gives its caller no CgResult, so a forward or
    adjoint solve that stops at max_iters raises DimError instead of
    returning a wrong x or wrong gradients."""
    import torch

    from diploma_research_amd import DimError
    from diploma_research_amd.autograd import solve

    systems = _systems([30, 41], 0.5, seed=41)
    A = Csr5.from_csr_batched(systems, sigma=8)
    vals = torch.from_numpy(np.concatenate([v for *_, v in systems])).cuda().requires_grad_(True)
    b = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, 900 + 1681)).cuda().requires_grad_(True)
    with pytest.raises(DimError, match="forward solve"):
        solve(A, vals, b, rtol=1e-12, max_iters=3)
    x = solve(A, vals, b, rtol=1e-12, max_iters=20000)
    # the custom Function's backward node is its ctx: cap the adjoint solve
    x.grad_fn.max_iters = 3
    with pytest.raises(DimError, match="adjoint solve"):
        (0.5 * (x * x).sum()).backward()


@pytest.mark.parametrize("persistent", [False, True])
def test_batched_chunks_mixing_stopped_and_running_systems(cuda, persistent):
    """Chunks of several tiles whose tiles belong to systems that stop early
    (tiny ones) and to one slow, large system placed between them: the
    product skips a chunk only when all of its systems have stopped, and the
    carry fix-up may then leave stale sums in stopped systems' rows of q,
    which no later pass reads. Every system, the large one above all, must
    still match the oracle's CG (iterations as _iters_close allows, solution
    within the tolerance's reach)."""
    rng = np.random.default_rng(77)
    sides = [int(s) for s in rng.integers(4, 12, size=60)]
    sides = sides[:30] + [700] + sides[30:]
    systems = _systems(sides, 0.5, seed=700)
    A = Csr5.from_csr_batched(systems, sigma=4)
    total_tiles = A.tile_count
    assert total_tiles > 8192, total_tiles  # more tiles than resident waves: chunks of several tiles
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    X, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=20000,
                                        persistent=persistent)
    Xh = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=20000)
        assert r.converged and _iters_close(r.iterations, oc["iterations"]), (n, r.iterations, oc["iterations"])
        err = np.linalg.norm(Xh[off:off + n] - oc["x"]) / max(np.linalg.norm(oc["x"]), 1e-300)
        assert err < 1e-7, (n, err)
        off += n


def _arrow_system(n, seed):
    """SPD arrow matrix: diagonal plus a full first row and column, so that
    row 0 spans many tiles (and chunks)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, n)
    u[0] = 0.0
    d = 2.0 + np.abs(u) * 2.0
    d[0] = 2.0 + float(np.sum(np.abs(u))) * 2.0
    rows, cols, vals = [], [], []
    for j in range(n):
        rows.append(0), cols.append(j), vals.append(d[0] if j == 0 else u[j])
    for i in range(1, n):
        rows += [i, i]
        cols += [0, i]
        vals += [u[i], d[i]]
    rows, cols, vals = np.array(rows), np.array(cols), np.array(vals)
    rp = np.zeros(n + 1, np.uint32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.uint32)
    order = np.lexsort((cols, rows))
    return n, rp, cols[order].astype(np.uint32), vals[order]


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("sigma", [4, 16])
def test_batched_rows_crossing_chunks(cuda, sigma, persistent):
    """Systems whose first row holds thousands of elements, so that rows
    cross several chunk boundaries and the product runs its carry fix-up
    (the separate kernel, or the persistent solve's fix-up phase): every
    system against the oracle's CG."""
    systems = [_arrow_system(n, seed=n) for n in (3000, 1700, 2500)] + _systems([20, 35], 0.5, seed=8)
    A = Csr5.from_csr_batched(systems, sigma=sigma)
    rng = np.random.default_rng(11)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    X, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=5000,
                                        persistent=persistent)
    X = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=5000)
        assert r.converged and _iters_close(r.iterations, oc["iterations"]), (n, r.iterations, oc["iterations"])
        np.testing.assert_allclose(X[off: off + n], oc["x"], rtol=1e-7, atol=1e-9)
        off += n
    # the batch's product over its chunk table, rows crossing chunks included
    x = rng.uniform(-1, 1, sum(n for n, *_ in systems))
    y = A.spmv(to_dev(x)).cpu().numpy()
    off = 0
    for n, rp, ci, v in systems:
        xs = x[off: off + n].copy()
        assert_spmv_close(y[off: off + n], O.csr_spmv(rp, ci, v, xs), rp, ci, v, xs)
        off += n


@pytest.mark.parametrize("sigma", [4, 5, 6, 7, 8])
def test_batched_column_codes_per_system(cuda, sigma):
    """Stencil systems of different sides keep column codes, each with its own
    table of diagonals (col - row = -s, -1, 0, 1, s): the batch's products,
    single and multi-RHS, against the oracle per system, and the solve."""
    import torch

    sides = [13, 40, 7, 29]
    systems = _systems(sides, 0.5, seed=60 + sigma)
    A = Csr5.from_csr_batched(systems, sigma=sigma)
    assert A.info.column_codes == 1
    N = sum(n for n, *_ in systems)
    rng = np.random.default_rng(sigma)
    X = rng.uniform(-1, 1, (N, 3))
    Y = A.spmm(to_dev(X)).cpu().numpy()
    y = A.spmv(to_dev(X[:, 0].copy())).cpu().numpy()
    off = 0
    for n, rp, ci, v in systems:
        for j in range(3):
            xs = X[off: off + n, j].copy()
            ref = O.csr_spmv(rp, ci, v, xs)
            assert_spmv_close(Y[off: off + n, j], ref, rp, ci, v, xs)
            if j == 0:
                assert_spmv_close(y[off: off + n], ref, rp, ci, v, xs)
        off += n
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    Xs, res = conjugate_gradient_batched(A, to_dev(np.concatenate(bs)), rtol=1e-10, max_iters=5000)
    Xs = Xs.cpu().numpy()
    off = 0
    for (n, rp, ci, v), b, r in zip(systems, bs, res):
        oc = O.Csr5(n, n, rp, ci, v).cg(b, rtol=1e-10, maxit=5000)
        assert r.converged and _iters_close(r.iterations, oc["iterations"])
        np.testing.assert_allclose(Xs[off: off + n], oc["x"], rtol=1e-7, atol=1e-8)
        off += n
    # the values can still be replaced (the codes keep the pattern)
    vals = torch.from_numpy(np.concatenate([v for *_, v in systems]) * 3.0).cuda()
    A.set_values(vals)
    np.testing.assert_allclose(A.spmv(to_dev(X[:, 0].copy())).cpu().numpy(), 3 * y, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("persistent", [False, True])
def test_batched_iteration_cap_and_reproducibility(cuda, persistent):
    """The solve stops every system at max_iters with the oracle's iterate of
    that many steps, and two solves are bitwise equal."""
    import torch

    systems = _systems([30, 45, 12], 0.5, seed=70)
    A = Csr5.from_csr_batched(systems, sigma=0)
    rng = np.random.default_rng(12)
    bs = [rng.uniform(-1, 1, n) for n, *_ in systems]
    b = to_dev(np.concatenate(bs))
    X, res = conjugate_gradient_batched(A, b, rtol=1e-14, max_iters=17, persistent=persistent)
    Xh = X.cpu().numpy()
    off = 0
    for (n, rp, ci, v), bb, r in zip(systems, bs, res):
        oc = O.Csr5(n, n, rp, ci, v).cg(bb, rtol=1e-14, maxit=17)
        assert r.converged == oc["converged"]
        if not r.converged:
            assert r.iterations == 17
        np.testing.assert_allclose(Xh[off: off + n], oc["x"], rtol=1e-9, atol=1e-11)
        off += n
    X1, _ = conjugate_gradient_batched(A, b, rtol=1e-12, max_iters=5000, persistent=persistent)
    X2, _ = conjugate_gradient_batched(A, b, rtol=1e-12, max_iters=5000, persistent=persistent)
    assert torch.equal(X1, X2)


@pytest.mark.parametrize("sigma", [0, 4])
def test_persistent_solve_equals_launch_sequence(cuda, sigma):
    """Both forms use the same chunk table, the same per-chunk p.q partials
    and the same 4096-row vector items reduced in the same order, so their
    iterates, statuses and iteration counts agree bit for bit; A^T with the
    persistent form is refused."""
    import torch

    systems = [_arrow_system(900, seed=5)] + _systems([9, 64, 31, 50], 0.5, seed=90)
    A = Csr5.from_csr_batched(systems, sigma=sigma)
    rng = np.random.default_rng(13)
    b = to_dev(rng.uniform(-1, 1, sum(n for n, *_ in systems)))
    X1, r1 = conjugate_gradient_batched(A, b, rtol=1e-11, max_iters=5000)
    X2, r2 = conjugate_gradient_batched(A, b, rtol=1e-11, max_iters=5000, persistent=True)
    assert torch.equal(X1, X2)
    assert [(r.iterations, r.converged, r.relres) for r in r1] == [(r.iterations, r.converged, r.relres) for r in r2]
    with pytest.raises(ValueError, match="persistent"):
        conjugate_gradient_batched(A, b, transpose=True, persistent=True)
