// Conjugate-gradient solvers on the CSR5 product.
//
// Textbook CG (thesis.tex:2133-2160) in the step order of the reference's
// single-process driver (benchmarks/conjugate_gradient/dim_sp/main.cpp:
// 102-127): As = A*s; alpha = r.r / s.As; x += alpha*s; r -= alpha*As;
// r.r_new; stop test; beta = r.r_new / r.r; s = r + beta*s.
//
// One iteration is four kernels with every scalar kept on the device:
//   1. CSR5 tiles  q = A p, with p[row]*q[row] folded into the epilogue;
//   2. carry fix-up of rows that cross tiles, which also reduces the p.q
//      partials of 256 tiles into one partial per block;
//   3. x += alpha p, r -= alpha q and the r.r partials (alpha from the p.q
//      partials, reduced redundantly and in a fixed order by every block);
//   4. p = r + beta p, the stop test and the iteration counter.
// There are no atomics, so iterates are bitwise reproducible. The r.r
// partials are double-buffered by iteration parity, and a batch of an even
// number of iterations is captured once in a hipGraph and replayed; every
// kernel returns at once when the device-side `done` flag is set, so the
// host only looks at the state between batches.
//
// k right-hand sides share one SpMM per iteration (independent CG per
// column). The FP32-refined mode iterates in FP32 (values, vectors; dot
// products accumulated in FP64) and replaces the residual in FP64 between
// batches: x += x_corr, r = b - A x, both in FP64, and the stop test uses that
// FP64 residual (reliable updates).
#include "handle.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace dimgpu {

namespace {

constexpr int kVecThreads = 1024;
constexpr int kMaxVecBlocks = 256;

struct CgState {
    int done; // 0 running, 1 converged, 2 breakdown, 3 iteration cap
    int iterations;
    int max_iters;
    int k;
    int col_done[kMaxK]; // 1 converged, 2 breakdown
    int col_iters[kMaxK];
    double tol2[kMaxK];  // (rtol * ||b_j||)^2
    double bnorm[kMaxK];
    double rr[kMaxK];     // r.r after the last completed iteration
    double rr_true[kMaxK]; // FP64 residual norm^2 at the last residual replacement
    double alpha[kMaxK];   // this iteration's step, from k_cg_update_r to k_cg_update_xp
    int broke;             // a column broke down in this iteration
    // `done` as k_cg_update_r saw it. k_cg_update_xp tests this copy: its own
    // block (0, 0) sets `done` at its end, and a block of the same launch
    // that starts after that must still apply this iteration's x update.
    int done_prev;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// Sum of parts[i*k + j], i < n, computed by wave 0 in a fixed order and
// broadcast through `out` (LDS). Every block obtains the same bits.
template<int KMAX>
__device__ void reduce_parts(const double* parts, int n, int k, double* out) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int i = lane; i < n; i += 64)
                s += parts[static_cast<size_t>(i) * k + j];
            s = wave_sum(s);
            if (lane == 0)
                out[j] = s;
        }
    }
    __syncthreads();
}

// Multi-RHS blocks live column-major inside the solver: element (i, j) of an
// n x k block is at j*n + i, so every column is a plain vector. Vector
// kernels run on a (blocks, k) grid with blockIdx.y the column; a block's
// partial sum for its column goes to parts[blockIdx.x*k + j] (fixed order).
__device__ void block_store_sum(double acc, double* parts) {
    __shared__ double red[kVecThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    acc = wave_sum(acc);
    if (lane == 0)
        red[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i)
            s += red[i];
        parts[static_cast<size_t>(blockIdx.x) * gridDim.y + blockIdx.y] = s;
    }
    __syncthreads();
}

// sum_i parts[i*k + j], i < n, in a fixed order (thread t sums i = t,
// t + blockDim.x, ...; then waves in order), broadcast through `out` (LDS):
// every block obtains the same bits. All threads load, so a few thousand
// partials cost one memory round trip.
__device__ void reduce_col(const double* parts, int n, int k, int j, double* out) {
    __shared__ double red[kVecThreads / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        s += parts[static_cast<size_t>(i) * k + j];
    s = wave_sum(s);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0)
        red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < nw; ++i)
            t += red[i];
        *out = t;
    }
    __syncthreads();
}

#define DIM_COL_LOOP(i)                                                                                               \
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;                               \
         i += static_cast<size_t>(gridDim.x) * blockDim.x)

// r = b - q (q = A x0), p = r; partials of r.r and b.b.
template<typename V>
__global__ __launch_bounds__(kVecThreads) void k_cg_init(size_t n, const double* b, const double* q, V* r, V* p,
                                                         double* rr_part, double* bb_part) {
    const size_t o = blockIdx.y * n;
    double acc = 0.0, bbs = 0.0;
    DIM_COL_LOOP(i) {
        const double bi = b[o + i];
        const V ri = static_cast<V>(bi - q[o + i]);
        r[o + i] = ri;
        p[o + i] = ri;
        acc += static_cast<double>(ri) * static_cast<double>(ri);
        bbs += bi * bi;
    }
    block_store_sum(acc, rr_part);
    block_store_sum(bbs, bb_part);
}

// Tolerances and the initial stop test (one block).
template<int KMAX>
__global__ void k_cg_init_state(CgState* st, int nparts, const double* rr_part, const double* bb_part, double rtol,
                                double* hist) {
    __shared__ double rr[KMAX], bb[KMAX];
    const int k = st->k;
    reduce_parts<KMAX>(rr_part, nparts, k, rr);
    reduce_parts<KMAX>(bb_part, nparts, k, bb);
    if (threadIdx.x == 0) {
        int all = 1;
        for (int j = 0; j < k; ++j) {
            const double bn = std::sqrt(bb[j]);
            st->bnorm[j] = bn;
            const double scale = bn > 0.0 ? bn : 1.0;
            st->tol2[j] = (rtol * scale) * (rtol * scale);
            st->rr[j] = rr[j];
            st->rr_true[j] = rr[j];
            st->col_done[j] = rr[j] <= st->tol2[j] ? 1 : 0;
            st->col_iters[j] = 0;
            all &= st->col_done[j] != 0;
            if (hist)
                hist[j] = std::sqrt(rr[j]) / scale;
        }
        st->iterations = 0;
        st->done = all ? 1 : (st->max_iters == 0 ? 3 : 0);
    }
}

// alpha = r.r / p.q; r -= alpha q; r.r partials. The step goes to the state
// for k_cg_update_xp, which moves x with the same p that p.q used: x and p
// are then read and written in one pass (8 vector passes per iteration
// instead of 9).
template<typename V>
__global__ __launch_bounds__(kVecThreads) void k_cg_update_r(CgState* st, size_t n, V* r, const V* q,
                                                             const double* pq_part, int n_pq, const double* rr_in,
                                                             double* rr_out, int n_rr) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
        st->done_prev = st->done;
    if (st->done)
        return;
    const int j = blockIdx.y, k = gridDim.y;
    __shared__ double pq, rr, alpha_s;
    reduce_col(pq_part, n_pq, k, j, &pq);
    reduce_col(rr_in, n_rr, k, j, &rr);
    if (threadIdx.x == 0) {
        double a = rr / pq;
        const bool broke = !(pq > 0.0) || !std::isfinite(a);
        const int cd = st->col_done[j];
        if (cd || broke)
            a = 0.0;
        alpha_s = a;
        if (blockIdx.x == 0) {
            st->alpha[j] = a;
            if (broke && !cd) {
                st->col_done[j] = 2;
                st->col_iters[j] = st->iterations;
                st->broke = 1;
            }
        }
    }
    __syncthreads();
    const V a = static_cast<V>(alpha_s);
    const size_t o = static_cast<size_t>(j) * n;
    double acc = 0.0;
    DIM_COL_LOOP(i) {
        const V ri = r[o + i] - a * q[o + i];
        r[o + i] = ri;
        acc += static_cast<double>(ri) * static_cast<double>(ri);
    }
    block_store_sum(acc, rr_out);
}

// x += alpha p; stop test; beta = r.r_new / r.r; p = r + beta p. Block (0, 0)
// keeps the iteration count and the per-column states.
template<typename V, int KMAX>
__global__ __launch_bounds__(kVecThreads) void k_cg_update_xp(CgState* st, size_t n, V* x, const V* r, V* p,
                                                              const double* rr_new_part, const double* rr_old_part,
                                                              int n_rr, double* hist, int hist_cap) {
    if (st->done_prev)
        return;
    const int j = blockIdx.y, k = gridDim.y;
    __shared__ double rn, ro;
    reduce_col(rr_new_part, n_rr, k, j, &rn);
    reduce_col(rr_old_part, n_rr, k, j, &ro);
    const V a = static_cast<V>(st->alpha[j]);
    // a column frozen earlier (converged or broken down) keeps r, so its
    // test repeats with the same numbers
    const bool conv = (rn <= st->tol2[j]) || (st->col_done[j] == 2);
    const size_t o = static_cast<size_t>(j) * n;
    if (!conv) {
        const V beta = static_cast<V>(rn / ro);
        DIM_COL_LOOP(i) {
            const V pi = p[o + i];
            x[o + i] = x[o + i] + a * pi;
            p[o + i] = r[o + i] + beta * pi;
        }
    } else {
        DIM_COL_LOOP(i) {
            x[o + i] = x[o + i] + a * p[o + i];
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        __shared__ double rn_all[KMAX];
        reduce_parts<KMAX>(rr_new_part, n_rr, k, rn_all);
        if (threadIdx.x == 0) {
            const int it = st->iterations + 1;
            st->iterations = it;
            int all = 1;
            for (int c = 0; c < k; ++c) {
                const int cc = (rn_all[c] <= st->tol2[c]) || (st->col_done[c] == 2);
                all &= cc;
                st->rr[c] = rn_all[c];
                if (cc && !st->col_done[c]) {
                    st->col_done[c] = 1;
                    st->col_iters[c] = it;
                }
                if (hist && it < hist_cap) {
                    const double scale = st->bnorm[c] > 0.0 ? st->bnorm[c] : 1.0;
                    hist[static_cast<size_t>(it) * k + c] = std::sqrt(rn_all[c]) / scale;
                }
            }
            if (st->broke)
                st->done = 2;
            else if (all)
                st->done = 1;
            else if (it >= st->max_iters)
                st->done = 3;
        }
    }
}

// Reliable update, part 1: x += x_corr (FP64), x_corr = 0.
__global__ __launch_bounds__(kVecThreads) void k_ru_fold(size_t nk, double* x, float* xc) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < nk;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        x[i] += static_cast<double>(xc[i]);
        xc[i] = 0.0f;
    }
}

// Reliable update, part 2: r64 = b - q64; r32 = float(r64); partials of the
// FP32 residual's r.r (iteration scalar) and of the FP64 residual (stop test).
// A column that the FP32 recursion has marked converged stopped updating p
// (k_cg_update_xp), so its p is stale: it restarts with p = r, which is the
// right direction whether k_ru_state then finds it converged (p unused) or
// resumes it. Columns still running keep p (the recurrence continues with
// the replaced residual).
__global__ __launch_bounds__(kVecThreads) void k_ru_residual(const CgState* st, size_t n, const double* b,
                                                             const double* q, float* r32, float* p32,
                                                             double* rr_part, double* rt_part) {
    const size_t o = blockIdx.y * n;
    const bool restart = st->col_done[blockIdx.y] == 1;
    double acc = 0.0, trs = 0.0;
    DIM_COL_LOOP(i) {
        const double rv = b[o + i] - q[o + i];
        const float rsv = static_cast<float>(rv);
        r32[o + i] = rsv;
        if (restart)
            p32[o + i] = rsv;
        acc += static_cast<double>(rsv) * static_cast<double>(rsv);
        trs += rv * rv;
    }
    block_store_sum(acc, rr_part);
    block_store_sum(trs, rt_part);
}

// Reliable update, part 3 (one block): the FP64 stop test decides; columns
// whose FP64 residual is not yet small enough are resumed.
template<int KMAX>
__global__ void k_ru_state(CgState* st, int nparts, const double* rr_part, const double* rt_part) {
    __shared__ double rr[KMAX], rt[KMAX];
    const int k = st->k;
    reduce_parts<KMAX>(rr_part, nparts, k, rr);
    reduce_parts<KMAX>(rt_part, nparts, k, rt);
    if (threadIdx.x == 0) {
        int all = 1;
        for (int j = 0; j < k; ++j) {
            st->rr[j] = rr[j];
            st->rr_true[j] = rt[j];
            if (st->col_done[j] != 2) {
                const int c = rt[j] <= st->tol2[j];
                if (c && st->col_done[j] != 1)
                    st->col_iters[j] = st->iterations;
                st->col_done[j] = c;
            }
            all &= st->col_done[j] != 0;
        }
        if (st->done != 2)
            st->done = all ? 1 : (st->iterations >= st->max_iters ? 3 : 0);
    }
}

// ||b - q||^2 partials (FP64) for the final true residual.
__global__ __launch_bounds__(kVecThreads) void k_resid_parts(size_t n, const double* b, const double* q,
                                                             double* parts) {
    const size_t o = blockIdx.y * n;
    double acc = 0.0;
    DIM_COL_LOOP(i) {
        const double d = b[o + i] - q[o + i];
        acc += d * d;
    }
    block_store_sum(acc, parts);
}

// Row-major n x k <-> the solver's column-major layout.
__global__ void k_rows_to_cols(size_t n, int k, const double* src, double* dst) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n * k;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[(e % k) * n + e / k] = src[e];
}
__global__ void k_cols_to_rows(size_t n, int k, const double* src, double* dst) {
    for (size_t e = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; e < n * k;
         e += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[e] = src[(e % k) * n + e / k];
}

template<typename V>
__global__ void k_zero(V* p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = V(0);
}

// ------------------------------------------------------------ host side --

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs)
            (void)hipFree(p);
    }
    template<typename T>
    hipError_t alloc(T*& p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess)
            ptrs.push_back(q);
        p = static_cast<T*>(q);
        return e;
    }
};

struct Graph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    void reset() {
        if (ge)
            (void)hipGraphExecDestroy(ge);
        if (g)
            (void)hipGraphDestroy(g);
        ge = nullptr;
        g = nullptr;
    }
    ~Graph() { reset(); }
};

struct PinnedState {
    CgState* h = nullptr;
    ~PinnedState() {
        if (h)
            (void)hipHostFree(h);
    }
};

uint32_t vec_blocks(size_t n) {
    return static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(kMaxVecBlocks, (n + 4 * kVecThreads - 1) /
                                                                                          (4 * kVecThreads))));
}

template<int KMAX>
struct Launcher {
    // one CG iteration in precision V (f32 picks the FP32 matrix values); the
    // n x k blocks are column-major (ldv = n)
    template<typename V>
    static void iteration(dim_csr5* A, CgState* st, size_t n, int k, V* x, V* r, V* p, V* q, double* rr_a,
                          double* rr_b, uint32_t nvb, double* hist, int hist_cap, hipStream_t s) {
        constexpr bool f32 = std::is_same<V, float>::value;
        uint32_t n_pq = 1;
        const double* pq_parts = A->s.dot_blk;
        (void)launch_spmv(A->f, A->s, f32, k, p, q, 0, A->f.tiles, 0, 0, A->f.n_empty, true, &st->done, s, &n_pq, n,
                          0, &pq_parts);
        const dim3 grid(nvb, static_cast<uint32_t>(k));
        hipLaunchKernelGGL((k_cg_update_r<V>), grid, dim3(kVecThreads), 0, s, st, n, r, static_cast<const V*>(q),
                           pq_parts, static_cast<int>(n_pq), rr_a, rr_b, static_cast<int>(nvb));
        hipLaunchKernelGGL((k_cg_update_xp<V, KMAX>), grid, dim3(kVecThreads), 0, s, st, n, x,
                           static_cast<const V*>(r), p, rr_b, rr_a, static_cast<int>(nvb), hist, hist_cap);
    }
};

// The device buffers of one solve: the state, the reduction partials (nvb *
// kMaxK each), the work vectors of the precision in use and, for k > 1, the
// column-major copies of B and X.
struct Work {
    CgState* st = nullptr;
    double *rr0 = nullptr, *rr1 = nullptr, *bb = nullptr, *rt = nullptr, *hist = nullptr;
    double *q64 = nullptr, *r64 = nullptr, *p64 = nullptr, *bc = nullptr, *xc = nullptr;
    float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr;
};

dim_status alloc_work(DevBuf& buf, size_t n, int k, uint32_t nvb, bool mixed, int hist_cap, Work& w) {
    DIM_HIP_TRY(buf.alloc(w.st, 1));
    DIM_HIP_TRY(buf.alloc(w.q64, n * k));
    DIM_HIP_TRY(buf.alloc(w.rr0, size_t{nvb} * kMaxK));
    DIM_HIP_TRY(buf.alloc(w.rr1, size_t{nvb} * kMaxK));
    DIM_HIP_TRY(buf.alloc(w.bb, size_t{nvb} * kMaxK));
    DIM_HIP_TRY(buf.alloc(w.rt, size_t{nvb} * kMaxK));
    if (hist_cap)
        DIM_HIP_TRY(buf.alloc(w.hist, static_cast<size_t>(hist_cap) * k));
    if (mixed) {
        DIM_HIP_TRY(buf.alloc(w.r32, n * k));
        DIM_HIP_TRY(buf.alloc(w.p32, n * k));
        DIM_HIP_TRY(buf.alloc(w.q32, n * k));
        DIM_HIP_TRY(buf.alloc(w.x32, n * k));
    } else {
        DIM_HIP_TRY(buf.alloc(w.r64, n * k));
        DIM_HIP_TRY(buf.alloc(w.p64, n * k));
    }
    if (k > 1) {
        DIM_HIP_TRY(buf.alloc(w.bc, n * k));
        DIM_HIP_TRY(buf.alloc(w.xc, n * k));
    }
    return DIM_SUCCESS;
}

// Bytes of the work buffers alloc_work() allocates.
size_t work_bytes(size_t n, int k, uint32_t nvb, bool mixed, int hist_cap) {
    const size_t nk = n * static_cast<size_t>(k);
    return sizeof(CgState) + 4 * size_t{nvb} * kMaxK * sizeof(double) + static_cast<size_t>(hist_cap) * k * 8 +
           nk * 8 /* q64 */ + nk * 16 /* r64, p64 or r32, p32, q32, x32 */ + (k > 1 ? nk * 16 : 0);
}

} // namespace

// A handle's last solve: its work buffers, the pinned copy of the state, the
// capture stream and the captured batch of iterations, reused by the next
// solve on the handle with the same rows, right-hand sides, precision and
// history length (and, for the graph, the same iteration kernels, batch
// length and x). A small solve then spends its time in the iterations, not
// in hipMalloc, hipHostMalloc and graph instantiation. Solves whose buffers
// exceed kCgCacheMaxBytes allocate per call instead, so that a handle never
// holds more than that between solves.
struct CgCache {
    DevBuf buf;
    Work w;
    size_t n = 0;
    int k = 0;
    bool mixed = false;
    int hist_cap = 0;
    PinnedState pinned;
    hipStream_t cap = nullptr;
    Graph graph;
    int g_kmax = 0;
    uint32_t g_check = 0;
    const void* g_x = nullptr;
    ~CgCache() {
        graph.reset();
        if (cap)
            (void)hipStreamDestroy(cap);
    }
};

void free_cg_cache(CgCache* c) {
    delete c;
}

namespace {

constexpr size_t kCgCacheMaxBytes = size_t{256} << 20;

// The handle's cache for a solve of this shape, made (and the old one freed)
// when the shape changed; nullptr when the buffers are larger than the cache
// keeps.
dim_status cg_cache_for(dim_csr5* A, size_t n, int k, uint32_t nvb, bool mixed, int hist_cap, CgCache** out) {
    *out = nullptr;
    if (work_bytes(n, k, nvb, mixed, hist_cap) > kCgCacheMaxBytes)
        return DIM_SUCCESS;
    CgCache* c = A->cg_cache;
    if (c && c->n == n && c->k == k && c->mixed == mixed && c->hist_cap == hist_cap) {
        *out = c;
        return DIM_SUCCESS;
    }
    free_cg_cache(c);
    A->cg_cache = nullptr;
    std::unique_ptr<CgCache> fresh(new CgCache);
    if (dim_status e = alloc_work(fresh->buf, n, k, nvb, mixed, hist_cap, fresh->w))
        return e;
    DIM_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&fresh->pinned.h), sizeof(CgState), hipHostMallocDefault));
    DIM_HIP_TRY(hipStreamCreateWithFlags(&fresh->cap, hipStreamNonBlocking));
    fresh->n = n;
    fresh->k = k;
    fresh->mixed = mixed;
    fresh->hist_cap = hist_cap;
    A->cg_cache = fresh.release();
    *out = A->cg_cache;
    return DIM_SUCCESS;
}

// fixed == nullptr: the solve allocates its buffers, captures a batch of
// iterations in a hipGraph and polls a pinned copy of the state. Otherwise its
// buffers are the ones `fixed` points into and it allocates nothing: no
// device, pinned or graph memory, no events (kernels are launched directly,
// the state is copied to the host stack, the time is host wall time).
template<int KMAX>
dim_status solve(dim_csr5* A, int k, const double* b_in, double* x_out, const dim_cg_opts& o, dim_cg_result* res,
                 hipStream_t stream, const Work* fixed = nullptr) {
    const size_t n = A->f.rows;
    const bool mixed = o.precision == DIM_CG_F32_REFINED;
    if (mixed && !A->f.vals32)
        return fail(DIM_INVALID_ARGUMENT, "FP32-refined CG needs a matrix built with DIM_CSR5_KEEP_F32");
    const bool no_alloc = fixed != nullptr;
    const uint32_t nvb = vec_blocks(n);
    const dim3 vgrid(nvb, static_cast<uint32_t>(k));
    const uint32_t check = o.check_every ? std::max<uint32_t>(2, (o.check_every + 1) & ~1u) : 16;
    const double delta = o.refine_delta > 0.0 ? o.refine_delta : 0.1;
    const int hist_cap = o.history ? static_cast<int>(o.max_iters) + 1 : 0;

    DevBuf buf;
    Work w;
    CgCache* cc = nullptr; // the handle's buffers, when this solve's fit them
    if (no_alloc) {
        w = *fixed;
    } else {
        if (dim_status e = cg_cache_for(A, n, k, nvb, mixed, hist_cap, &cc))
            return e;
        if (cc)
            w = cc->w;
        else if (dim_status e = alloc_work(buf, n, k, nvb, mixed, hist_cap, w))
            return e;
    }
    CgState* st = w.st;
    double *q64 = w.q64, *r64 = w.r64, *p64 = w.p64, *rr0 = w.rr0, *rr1 = w.rr1, *bb = w.bb, *rt = w.rt,
           *hist = w.hist;
    float *r32 = w.r32, *p32 = w.p32, *q32 = w.q32, *x32 = w.x32;
    const double* b = k > 1 ? w.bc : b_in;
    double* x = k > 1 ? w.xc : x_out;
    PinnedState pinned;
    CgState stack_state{};
    CgState* hs = &stack_state;
    if (cc) {
        hs = cc->pinned.h;
    } else if (!no_alloc) {
        DIM_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&pinned.h), sizeof(CgState), hipHostMallocDefault));
        hs = pinned.h;
    }

    auto enqueue_batch = [&](hipStream_t s) {
        for (uint32_t it = 0; it < check; it += 2) {
            if (mixed) {
                Launcher<KMAX>::template iteration<float>(A, st, n, k, x32, r32, p32, q32, rr0, rr1, nvb, hist,
                                                          hist_cap, s);
                Launcher<KMAX>::template iteration<float>(A, st, n, k, x32, r32, p32, q32, rr1, rr0, nvb, hist,
                                                          hist_cap, s);
            } else {
                Launcher<KMAX>::template iteration<double>(A, st, n, k, x, r64, p64, q64, rr0, rr1, nvb, hist,
                                                           hist_cap, s);
                Launcher<KMAX>::template iteration<double>(A, st, n, k, x, r64, p64, q64, rr1, rr0, nvb, hist,
                                                           hist_cap, s);
            }
        }
    };
    // capture `check` iterations (even) once, or reuse the handle's capture
    // of the same iterations
    Graph graph;
    hipGraphExec_t exec = nullptr;
    if (cc && cc->graph.ge && cc->g_kmax == KMAX && cc->g_check == check && cc->g_x == x) {
        exec = cc->graph.ge;
    } else if (!no_alloc) {
        Graph& gr = cc ? cc->graph : graph;
        gr.reset();
        hipStream_t cap = cc ? cc->cap : nullptr;
        if (!cc)
            DIM_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            enqueue_batch(cap);
            ok = hipStreamEndCapture(cap, &gr.g) == hipSuccess && gr.g &&
                 hipGraphInstantiate(&gr.ge, gr.g, nullptr, nullptr, 0) == hipSuccess;
            (void)hipGetLastError();
        }
        if (!cc)
            (void)hipStreamDestroy(cap);
        if (ok) {
            exec = gr.ge;
            if (cc) {
                cc->g_kmax = KMAX;
                cc->g_check = check;
                cc->g_x = x;
            }
        } else {
            gr.reset();
        }
    }
    const bool use_graph = exec != nullptr;

    hipEvent_t e0 = nullptr, e1 = nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    if (!no_alloc) {
        DIM_HIP_TRY(hipEventCreate(&e0));
        DIM_HIP_TRY(hipEventCreate(&e1));
        DIM_HIP_TRY(hipEventRecord(e0, stream));
    }

    const uint32_t tgrid = 1024;
    if (k > 1) {
        hipLaunchKernelGGL(k_rows_to_cols, dim3(tgrid), dim3(256), 0, stream, n, k, b_in, w.bc);
        hipLaunchKernelGGL(k_rows_to_cols, dim3(tgrid), dim3(256), 0, stream, n, k, static_cast<const double*>(x_out),
                           w.xc);
    }
    // Every buffer is written before it is read (the state by this copy, the
    // partials by whole-grid passes, x32 by k_zero), so memory left by an
    // earlier solve needs no clearing.
    CgState init{};
    init.max_iters = static_cast<int>(o.max_iters);
    init.k = k;
    DIM_HIP_TRY(hipMemcpyAsync(st, &init, sizeof(CgState), hipMemcpyHostToDevice, stream));
    // q = A x0 (FP64), r = b - q, p = r
    DIM_HIP_TRY(launch_spmv(A->f, A->s, false, k, x, q64, 0, A->f.tiles, 0, 0, A->f.n_empty, false, nullptr, stream,
                            nullptr, n));
    if (mixed) {
        hipLaunchKernelGGL((k_cg_init<float>), vgrid, dim3(kVecThreads), 0, stream, n, b, q64, r32, p32, rr0, bb);
        hipLaunchKernelGGL((k_zero<float>), dim3(nvb), dim3(kVecThreads), 0, stream, x32, n * k);
    } else {
        hipLaunchKernelGGL((k_cg_init<double>), vgrid, dim3(kVecThreads), 0, stream, n, b, q64, r64, p64, rr0, bb);
    }
    hipLaunchKernelGGL((k_cg_init_state<KMAX>), dim3(1), dim3(64), 0, stream, st, static_cast<int>(nvb), rr0, bb,
                       o.rtol, hist);
    DIM_HIP_TRY(hipGetLastError());

    uint32_t refinements = 0;
    std::array<double, kMaxK> rr_at_update{};
    bool first = true;
    dim_status status = DIM_SUCCESS;
    for (;;) {
        DIM_HIP_TRY(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
        DIM_HIP_TRY(hipStreamSynchronize(stream));
        if (first) {
            for (int j = 0; j < k; ++j)
                rr_at_update[j] = hs->rr_true[j];
            first = false;
        }
        if (mixed && hs->done != 2) {
            // residual replacement when a column's FP32 residual fell by delta
            // since the last one, when the FP32 residual says converged, or at
            // the iteration cap
            bool need = hs->done == 1 || hs->done == 3;
            for (int j = 0; j < k && !need; ++j)
                need = !hs->col_done[j] && hs->rr[j] < delta * delta * rr_at_update[j];
            if (need && hs->iterations > 0) {
                ++refinements;
                hipLaunchKernelGGL(k_ru_fold, dim3(nvb), dim3(kVecThreads), 0, stream, n * k, x, x32);
                DIM_HIP_TRY(launch_spmv(A->f, A->s, false, k, x, q64, 0, A->f.tiles, 0, 0, A->f.n_empty, false,
                                        nullptr, stream, nullptr, n));
                hipLaunchKernelGGL(k_ru_residual, vgrid, dim3(kVecThreads), 0, stream, static_cast<const CgState*>(st),
                                   n, b, static_cast<const double*>(q64), r32, p32, rr0, rt);
                hipLaunchKernelGGL((k_ru_state<KMAX>), dim3(1), dim3(64), 0, stream, st, static_cast<int>(nvb), rr0,
                                   rt);
                DIM_HIP_TRY(hipGetLastError());
                DIM_HIP_TRY(hipMemcpyAsync(hs, st, sizeof(CgState), hipMemcpyDeviceToHost, stream));
                DIM_HIP_TRY(hipStreamSynchronize(stream));
                for (int j = 0; j < k; ++j)
                    rr_at_update[j] = hs->rr[j];
            }
        }
        if (hs->done)
            break;
        if (use_graph) {
            DIM_HIP_TRY(hipGraphLaunch(exec, stream));
        } else {
            enqueue_batch(stream);
            DIM_HIP_TRY(hipGetLastError());
        }
    }
    // true residual ||b - A x|| in FP64, and the solution back in the
    // caller's row-major layout
    DIM_HIP_TRY(launch_spmv(A->f, A->s, false, k, x, q64, 0, A->f.tiles, 0, 0, A->f.n_empty, false, nullptr, stream,
                            nullptr, n));
    hipLaunchKernelGGL(k_resid_parts, vgrid, dim3(kVecThreads), 0, stream, n, b, static_cast<const double*>(q64), rt);
    if (k > 1)
        hipLaunchKernelGGL(k_cols_to_rows, dim3(tgrid), dim3(256), 0, stream, n, k,
                           static_cast<const double*>(w.xc), x_out);
    DIM_HIP_TRY(hipGetLastError());
    float ms = 0.f;
    if (!no_alloc) {
        DIM_HIP_TRY(hipEventRecord(e1, stream));
        DIM_HIP_TRY(hipEventSynchronize(e1));
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    // nvb <= kMaxVecBlocks partials per column
    std::array<double, size_t{kMaxVecBlocks} * kMaxK> parts;
    DIM_HIP_TRY(hipMemcpyAsync(parts.data(), rt, size_t{nvb} * k * sizeof(double), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipStreamSynchronize(stream));
    if (no_alloc)
        ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (hist_cap) {
        std::vector<double> hh(static_cast<size_t>(hist_cap) * k);
        DIM_HIP_TRY(hipMemcpy(hh.data(), hist, hh.size() * sizeof(double), hipMemcpyDeviceToHost));
        // opts.history holds column 0 (single right-hand side use)
        for (int i = 0; i < hist_cap; ++i)
            o.history[i] = hh[static_cast<size_t>(i) * k];
    }

    const CgState& S = *hs;
    for (int j = 0; j < k; ++j) {
        double tr = 0.0;
        for (uint32_t i = 0; i < nvb; ++i)
            tr += parts[size_t{i} * k + j];
        const double scale = S.bnorm[j] > 0.0 ? S.bnorm[j] : 1.0;
        dim_cg_result& R = res[j];
        R.converged = S.col_done[j] == 1;
        R.iterations = static_cast<uint32_t>(S.col_done[j] ? S.col_iters[j] : S.iterations);
        R.relres = std::sqrt(mixed ? S.rr_true[j] : S.rr[j]) / scale;
        R.relres_true = std::sqrt(tr) / scale;
        R.refinements = refinements;
        R.seconds = ms * 1e-3;
        if (S.col_done[j] == 2)
            status = DIM_BREAKDOWN;
        else if (!R.converged && status == DIM_SUCCESS)
            status = DIM_NOT_CONVERGED;
    }
    if (status == DIM_BREAKDOWN)
        set_error("CG breakdown: p'Ap <= 0 or not finite (matrix not SPD?)");
    else if (status == DIM_NOT_CONVERGED)
        set_error("CG reached max_iters above the tolerance");
    return status;
}

dim_status run(dim_csr5* A, int k, const double* b, double* x, const dim_cg_opts* opts, dim_cg_result* res,
               hipStream_t s) {
    if (k == 1)
        return solve<1>(A, k, b, x, *opts, res, s);
    if (k <= 2)
        return solve<2>(A, k, b, x, *opts, res, s);
    if (k <= 4)
        return solve<4>(A, k, b, x, *opts, res, s);
    if (k <= 8)
        return solve<8>(A, k, b, x, *opts, res, s);
    return solve<16>(A, k, b, x, *opts, res, s);
}

// ---- allocation-free solves (dim_cg_noalloc) -------------------------------
// One right-hand side of at most kNoAllocMaxRows rows. Layout of the array:
// the state, four partial arrays (rr0, rr1, bb, rt), then q64 and either r64,
// p64 (FP64) or r32, p32, q32, x32 (FP32-refined): 24 bytes per row either
// way, each piece starting on a 256-byte boundary.
constexpr size_t kNoAllocMaxRows = DIM_CG_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocAlign = 256;
constexpr size_t kNoAllocStateBytes = 1024;
constexpr size_t kNoAllocPartBytes = size_t{kMaxVecBlocks} * kMaxK * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocStateBytes + 4 * kNoAllocPartBytes + 24 * kNoAllocMaxRows + 8 * kNoAllocAlign;
static_assert(sizeof(CgState) <= kNoAllocStateBytes, "CgState must fit its slot");
static_assert(kNoAllocBytes == DIM_CG_NOALLOC_BYTES, "the header states the array's size");

// Each GPU holds its own copy from the library's load until the process
// ends. Host code must not use this name as an address: the address on each
// device comes from hipGetSymbolAddress (scratch_base).
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_base[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
dim_status scratch_base(char** out) {
    int dev = 0;
    DIM_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(DIM_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_base[dev]) {
        void* p = nullptr;
        DIM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_base[dev] = static_cast<char*>(p);
    }
    *out = g_scratch_base[dev];
    return DIM_SUCCESS;
}

// Carves the buffers of an n-row solve out of the array at `base`.
Work carve_noalloc(char* base, size_t n, bool mixed) {
    auto up = [](size_t b) { return (b + kNoAllocAlign - 1) / kNoAllocAlign * kNoAllocAlign; };
    Work w;
    size_t at = 0;
    w.st = reinterpret_cast<CgState*>(base);
    at += kNoAllocStateBytes;
    double** parts[4] = {&w.rr0, &w.rr1, &w.bb, &w.rt};
    for (double** pp : parts) {
        *pp = reinterpret_cast<double*>(base + at);
        at += kNoAllocPartBytes;
    }
    w.q64 = reinterpret_cast<double*>(base + at);
    at += up(n * sizeof(double));
    if (mixed) {
        float** vs[4] = {&w.r32, &w.p32, &w.q32, &w.x32};
        for (float** pv : vs) {
            *pv = reinterpret_cast<float*>(base + at);
            at += up(n * sizeof(float));
        }
    } else {
        w.r64 = reinterpret_cast<double*>(base + at);
        at += up(n * sizeof(double));
        w.p64 = reinterpret_cast<double*>(base + at);
    }
    return w;
}

dim_status check_cg_args(const dim_csr5* A, uint32_t k, const void* b, const void* x, const dim_cg_opts* opts,
                         const dim_cg_result* res) {
    if (!A)
        return fail(DIM_INVALID_ARGUMENT, "matrix handle is null");
    if (!opts || !res)
        return fail(DIM_INVALID_ARGUMENT, "opts and res are required");
    if (A->f.rows != A->f.cols)
        return fail(DIM_INVALID_ARGUMENT, "CG needs a square matrix");
    if (k < 1 || k > static_cast<uint32_t>(kMaxK))
        return fail(DIM_INVALID_ARGUMENT, "k must be in 1..16");
    if (A->f.rows && (!b || !x))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    if (!(opts->rtol >= 0.0))
        return fail(DIM_INVALID_ARGUMENT, "rtol must be >= 0");
    if (opts->precision != DIM_CG_F64 && opts->precision != DIM_CG_F32_REFINED)
        return fail(DIM_INVALID_ARGUMENT, "unknown precision");
    if (opts->history && k != 1)
        return fail(DIM_INVALID_ARGUMENT, "history is recorded for single right-hand sides only");
    if (A->n_systems != 1)
        return fail(DIM_INVALID_ARGUMENT, "batched matrices are solved with dim_cg_batched");
    return DIM_SUCCESS;
}

} // namespace
} // namespace dimgpu

using namespace dimgpu;

extern "C" {

dim_status dim_cg(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts, dim_cg_result* res,
                  void* stream) {
    if (dim_status st = check_cg_args(A, 1, b, x, opts, res))
        return st;
    return run(const_cast<dim_csr5*>(A), 1, b, x, opts, res, static_cast<hipStream_t>(stream));
}

dim_status dim_cg_host(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts, dim_cg_result* res,
                       void* stream) {
    if (dim_status st = check_cg_args(A, 1, b, x, opts, res))
        return st;
    auto* self = const_cast<dim_csr5*>(A);
    const auto s = static_cast<hipStream_t>(stream);
    char *db = nullptr, *dx = nullptr;
    if (dim_status st = stage_host_vectors(self, &db, &dx))
        return st;
    const size_t bytes = size_t{A->f.rows} * sizeof(double);
    DIM_HIP_TRY(hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s));
    DIM_HIP_TRY(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
    const dim_status st = run(self, 1, reinterpret_cast<const double*>(db), reinterpret_cast<double*>(dx), opts, res, s);
    if (st != DIM_SUCCESS && st != DIM_NOT_CONVERGED && st != DIM_BREAKDOWN)
        return st;
    DIM_HIP_TRY(hipMemcpyAsync(x, dx, bytes, hipMemcpyDeviceToHost, s));
    DIM_HIP_TRY(hipStreamSynchronize(s));
    return st;
}

dim_status dim_cg_noalloc(const dim_csr5* A, const double* b, double* x, const dim_cg_opts* opts,
                          dim_cg_result* res, void* stream) {
    if (dim_status st = check_cg_args(A, 1, b, x, opts, res))
        return st;
    if (opts->history)
        return fail(DIM_INVALID_ARGUMENT, "dim_cg_noalloc records no history (opts->history must be NULL)");
    if (A->f.rows > kNoAllocMaxRows)
        return fail(DIM_INVALID_ARGUMENT, "dim_cg_noalloc solves at most " + std::to_string(kNoAllocMaxRows) +
                                            " rows (the matrix has " + std::to_string(A->f.rows) + ")");
    int dev = 0;
    DIM_HIP_TRY(hipGetDevice(&dev));
    if (dev != A->device)
        return fail(DIM_INVALID_ARGUMENT, "the matrix lives on device " + std::to_string(A->device) +
                                            ", which must be the current device");
    char* base = nullptr;
    if (dim_status st = scratch_base(&base))
        return st;
    const Work w = carve_noalloc(base, A->f.rows, opts->precision == DIM_CG_F32_REFINED);
    return solve<1>(const_cast<dim_csr5*>(A), 1, b, x, *opts, res, static_cast<hipStream_t>(stream), &w);
}

dim_status dim_cg_multi(const dim_csr5* A, uint32_t k, const double* B, double* X, const dim_cg_opts* opts,
                        dim_cg_result* res, void* stream) {
    if (dim_status st = check_cg_args(A, k, B, X, opts, res))
        return st;
    return run(const_cast<dim_csr5*>(A), static_cast<int>(k), B, X, opts, res, static_cast<hipStream_t>(stream));
}

} // extern "C"
