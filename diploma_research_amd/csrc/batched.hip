// Many independent SPD systems solved together: one block-diagonal CSR5
// matrix (every system padded to whole tiles, so no tile straddles two
// systems) and a CG whose scalars are kept per system.
//
// One iteration is five kernels on the whole batch: the CSR5 tile kernel and
// its carry fix-up (q = A p, or A^T p for the adjoint solve), the per-system
// dot p.q, the x/r update, and the p update with each system's own stop test.
// Vector kernels run one block per slice of one system, so every block reads
// only its own system's scalars; the partial sums of a system are reduced in
// a fixed order, so iterates are reproducible (apart from the adjoint product,
// which scatters with atomics). A batch of iterations is captured once in a
// hipGraph and replayed until every system has stopped.
#include "handle.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

namespace dimgpu {

namespace {

constexpr int kThreads = 1024;
constexpr uint32_t kRowsPerBlock = 4 * kThreads;

// ------------------------------------------------------------ building ---

struct SystemCsr {
    uint32_t rows;
    uint64_t nnz;
    const uint32_t* row_ptr;
    const uint32_t* col_idx;
    const double* vals;
};

std::string build_batched(const std::vector<SystemCsr>& sys, uint32_t sigma, HostFormat& out,
                          std::vector<uint32_t>& sys_row, std::vector<uint32_t>& sys_tile,
                          std::vector<uint64_t>& sys_nnz) {
    out = HostFormat{};
    out.sigma = sigma;
    sys_row.assign(1, 0);
    sys_tile.assign(1, 0);
    sys_nnz.assign(1, 0);
    uint64_t rows = 0, tiles = 0, tdo_total = 0;
    for (const SystemCsr& s : sys) {
        HostFormat f;
        std::string err = build_format(s.rows, s.rows, s.nnz, s.row_ptr, s.col_idx, s.vals, sigma, f);
        if (!err.empty())
            return err;
        const uint32_t R = static_cast<uint32_t>(rows);
        for (uint32_t t = 0; t < f.tiles; ++t) {
            const uint32_t tp = f.tile_ptr[t];
            out.tile_ptr.push_back(((tp & ~kDirty) + R) | (tp & kDirty));
            out.tile_flags.push_back(f.tile_flags[t]);
            out.tile_desc_offset_ptr.push_back(static_cast<uint32_t>(tdo_total + f.tile_desc_offset_ptr[t]));
        }
        out.bit_flag.insert(out.bit_flag.end(), f.bit_flag.begin(), f.bit_flag.end());
        const uint32_t ntdo = f.tile_desc_offset_ptr[f.tiles];
        out.tile_desc_offset.insert(out.tile_desc_offset.end(), f.tile_desc_offset.begin(),
                                    f.tile_desc_offset.begin() + ntdo);
        tdo_total += ntdo;
        for (uint32_t& c : f.col_idx)
            c += R; // padding columns point at the system's first row
        out.col_idx.insert(out.col_idx.end(), f.col_idx.begin(), f.col_idx.end());
        out.vals.insert(out.vals.end(), f.vals.begin(), f.vals.end());
        for (uint32_t e : f.empty_rows)
            out.empty_rows.push_back(e + R);
        out.any_dirty |= f.any_dirty;
        rows += s.rows;
        tiles += f.tiles;
        if (rows >= kDirty || tiles >= kDirty)
            return "batch too large";
        sys_row.push_back(static_cast<uint32_t>(rows));
        sys_tile.push_back(static_cast<uint32_t>(tiles));
        sys_nnz.push_back(sys_nnz.back() + s.nnz);
    }
    out.rows = out.cols = static_cast<uint32_t>(rows);
    out.tiles = static_cast<uint32_t>(tiles);
    out.tile_ptr.push_back(out.rows);
    out.tile_flags.push_back(kStartsRow);
    out.tile_desc_offset_ptr.push_back(static_cast<uint32_t>(tdo_total));
    if (out.tile_desc_offset.empty())
        out.tile_desc_offset.push_back(0);
    out.padded = out.vals.size();
    // the kernels mask padding by position only in the last tile
    const size_t tile = size_t{kOmega} * sigma;
    const uint64_t last_nnz = sys.empty() ? 0 : sys.back().nnz;
    const uint64_t last_pad = sys.empty() ? 0 : (sys_tile[sys.size()] - sys_tile[sys.size() - 1]) * tile;
    out.nnz = out.padded - last_pad + last_nnz;
    return {};
}

dim_status finish_batched(HostFormat& f, bool keep_f32, const std::vector<uint32_t>& sys_row,
                          const std::vector<uint32_t>& sys_tile, const std::vector<uint64_t>& sys_nnz, dim_csr5** out) {
    dim_csr5* h = nullptr;
    if (dim_status st = upload_format(f, keep_f32, &h))
        return st;
    h->n_systems = static_cast<uint32_t>(sys_row.size() - 1);
    h->sys_row = sys_row;
    h->sys_tile = sys_tile;
    h->sys_nnz_csr = sys_nnz;
    h->nnz_logical = sys_nnz.back();
    *out = h;
    return DIM_SUCCESS;
}

// ------------------------------------------------------------- kernels ---

struct Tables {
    const uint32_t* blk_sys; // block -> system
    const uint32_t* blk_r0;  // first row of the block
    const uint32_t* blk_r1;  // end row of the block
    const uint32_t* sys_blk; // first block of each system (n_sys + 1)
};

struct BState {
    int done; // 0 running, 1 all systems stopped, 3 iteration cap
    int iterations;
    int max_iters;
    int n_sys;
    // `done` as k_b_update_r saw it. k_b_update_xp tests this copy: its own
    // block 0 sets `done` at its end, and a block of the same launch that
    // starts after that must still apply this iteration's x update.
    int done_prev;
};

// per-system arrays (n_sys entries each)
struct SysArrays {
    int* sys_done;   // 1 converged, 2 breakdown
    int* sys_iters;
    double* tol2;
    double* bnorm;
    double* rr;
    double* alpha;   // this iteration's step (0 for a stopped system)
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// sum of parts[b0..b1) by wave 0, fixed order; result broadcast through *out
__device__ void range_sum(const double* parts, uint32_t b0, uint32_t b1, double* out) {
    if (threadIdx.x < 64) {
        double s = 0.0;
        for (uint32_t i = b0 + threadIdx.x; i < b1; i += 64)
            s += parts[i];
        s = wave_sum(s);
        if (threadIdx.x == 0)
            *out = s;
    }
    __syncthreads();
}

__device__ void block_store_sum(double acc, double* parts) {
    __shared__ double red[kThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double s = wave_sum(acc);
    if (lane == 0)
        red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i)
            t += red[i];
        parts[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kThreads) void k_b_init(Tables tb, const double* b, const double* q, double* r, double* p,
                                                     double* rr_part, double* bb_part) {
    const uint32_t r0 = tb.blk_r0[blockIdx.x], r1 = tb.blk_r1[blockIdx.x];
    double acc = 0.0, bb = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const double ri = b[i] - q[i];
        r[i] = ri;
        p[i] = ri;
        acc += ri * ri;
        bb += b[i] * b[i];
    }
    block_store_sum(acc, rr_part);
    __syncthreads();
    block_store_sum(bb, bb_part);
}

__global__ void k_b_init_state(BState* st, SysArrays sa, Tables tb, const double* rr_part, const double* bb_part,
                               double rtol) {
    __shared__ double rr, bb;
    int all = 1;
    for (int s = 0; s < st->n_sys; ++s) {
        range_sum(rr_part, tb.sys_blk[s], tb.sys_blk[s + 1], &rr);
        range_sum(bb_part, tb.sys_blk[s], tb.sys_blk[s + 1], &bb);
        if (threadIdx.x == 0) {
            const double bn = std::sqrt(bb);
            const double scale = bn > 0.0 ? bn : 1.0;
            sa.bnorm[s] = bn;
            sa.tol2[s] = (rtol * scale) * (rtol * scale);
            sa.rr[s] = rr;
            sa.sys_done[s] = rr <= sa.tol2[s] ? 1 : 0;
            sa.sys_iters[s] = 0;
            all &= sa.sys_done[s] != 0;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st->iterations = 0;
        st->done = all ? 1 : (st->max_iters == 0 ? 3 : 0);
    }
}

// Systems that have stopped skip their vector work: their alpha is 0, so x
// and r stay as they are, and their partials are not used.
__global__ __launch_bounds__(kThreads) void k_b_dot(const BState* st, SysArrays sa, Tables tb, const double* p,
                                                    const double* q, double* pq_part) {
    if (st->done)
        return;
    double acc = 0.0;
    if (!sa.sys_done[tb.blk_sys[blockIdx.x]]) {
        const uint32_t r0 = tb.blk_r0[blockIdx.x], r1 = tb.blk_r1[blockIdx.x];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x)
            acc += p[i] * q[i];
    }
    block_store_sum(acc, pq_part);
}

// alpha = r.r / p.q per system; r -= alpha q; r.r partials. The step goes
// to sa.alpha for k_b_update_xp, which moves x with the same p that p.q
// used, in its pass over p (x is then read and written once per iteration).
__global__ __launch_bounds__(kThreads) void k_b_update_r(BState* st, SysArrays sa, Tables tb, double* r,
                                                         const double* q, const double* pq_part, const double* rr_in,
                                                         double* rr_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        st->done_prev = st->done;
    if (st->done)
        return;
    const uint32_t s = tb.blk_sys[blockIdx.x];
    __shared__ double pq, rr, alpha;
    __shared__ int skip;
    range_sum(pq_part, tb.sys_blk[s], tb.sys_blk[s + 1], &pq);
    range_sum(rr_in, tb.sys_blk[s], tb.sys_blk[s + 1], &rr);
    if (threadIdx.x == 0) {
        double a = rr / pq;
        const bool broke = !(pq > 0.0) || !std::isfinite(a);
        // every block of the system decides alike: a stopped system (or one
        // breaking down now, which its first block records) keeps x and r
        const bool stop = sa.sys_done[s] || broke;
        if (stop)
            a = 0.0;
        alpha = a;
        skip = stop;
        if (blockIdx.x == tb.sys_blk[s]) {
            sa.alpha[s] = a;
            if (broke && !sa.sys_done[s]) {
                sa.sys_done[s] = 2;
                sa.sys_iters[s] = st->iterations;
            }
        }
    }
    __syncthreads();
    const double al = alpha;
    double acc = 0.0;
    if (!skip) {
        const uint32_t r0 = tb.blk_r0[blockIdx.x], r1 = tb.blk_r1[blockIdx.x];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            const double ri = r[i] - al * q[i];
            r[i] = ri;
            acc += ri * ri;
        }
    }
    block_store_sum(acc, rr_out);
}

// x += alpha p and p = r + beta p in one pass (p read once); stop tests.
__global__ __launch_bounds__(kThreads) void k_b_update_xp(BState* st, SysArrays sa, Tables tb, double* x,
                                                          const double* r, double* p, const double* rr_new_part,
                                                          const double* rr_old_part) {
    if (st->done_prev)
        return;
    const uint32_t s = tb.blk_sys[blockIdx.x];
    __shared__ double rn, ro;
    range_sum(rr_new_part, tb.sys_blk[s], tb.sys_blk[s + 1], &rn);
    range_sum(rr_old_part, tb.sys_blk[s], tb.sys_blk[s + 1], &ro);
    const double al = sa.alpha[s];
    // a stopped system left its r.r partials at 0 and keeps p
    const bool conv = sa.sys_done[s] || rn <= sa.tol2[s];
    const double beta = rn / ro;
    if (al != 0.0 || !conv) {
        const uint32_t r0 = tb.blk_r0[blockIdx.x], r1 = tb.blk_r1[blockIdx.x];
        for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
            const double pi = p[i];
            x[i] = x[i] + al * pi;
            if (!conv)
                p[i] = r[i] + beta * pi;
        }
    }
    if (blockIdx.x != 0)
        return;
    // Block 0 keeps the batch state. Wave w tests systems w, w + 16, ...; each
    // sum runs in range_sum's order (lane l adds entries l, l + 64, ..., then
    // the butterfly), so the bits do not depend on which wave sums.
    __shared__ int wave_all[kThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int it = st->iterations + 1;
    int all = 1;
    for (int q = w; q < st->n_sys; q += nw) {
        double acc = 0.0;
        for (uint32_t i = tb.sys_blk[q] + lane; i < tb.sys_blk[q + 1]; i += 64)
            acc += rr_new_part[i];
        const double rs = wave_sum(acc);
        if (lane == 0) {
            // a stopped system keeps the r.r of its last step (its r no
            // longer changes)
            if (!sa.sys_done[q])
                sa.rr[q] = rs;
            if (!sa.sys_done[q] && rs <= sa.tol2[q]) {
                sa.sys_done[q] = 1;
                sa.sys_iters[q] = it;
            }
            all &= sa.sys_done[q] != 0;
        }
    }
    if (lane == 0)
        wave_all[w] = all;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < nw; ++i)
            all &= wave_all[i];
        st->iterations = it;
        if (all)
            st->done = 1;
        else if (it >= st->max_iters)
            st->done = 3;
    }
}

__global__ __launch_bounds__(kThreads) void k_b_resid(Tables tb, const double* b, const double* q, double* parts) {
    const uint32_t r0 = tb.blk_r0[blockIdx.x], r1 = tb.blk_r1[blockIdx.x];
    double acc = 0.0;
    for (uint32_t i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const double d = b[i] - q[i];
        acc += d * d;
    }
    block_store_sum(acc, parts);
}

// ------------------------------------------------------------ host side --

struct DevMem {
    std::vector<void*> ptrs;
    ~DevMem() {
        for (void* p : ptrs)
            (void)hipFree(p);
    }
    template<typename T>
    hipError_t alloc(T*& p, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess)
            ptrs.push_back(q);
        p = static_cast<T*>(q);
        return e;
    }
    template<typename T>
    hipError_t upload(T*& p, const std::vector<T>& v) {
        hipError_t e = alloc(p, v.size());
        if (e == hipSuccess && !v.empty())
            e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
        return e;
    }
};

// tile_sys, sys_done (may be null): chunks of stopped systems are skipped.
hipError_t product(dim_csr5* A, bool transpose, const double* x, double* y, const int* done, hipStream_t s,
                   const uint32_t* tile_sys = nullptr, const int* sys_done = nullptr) {
    if (transpose)
        return launch_spmv_transpose(A->f, false, x, y, s, done);
    return launch_spmv(A->f, A->s, false, 1, x, y, 0, A->f.tiles, 0, 0, A->f.n_empty, false, done, s, nullptr, 0, 0,
                       nullptr, tile_sys, sys_done);
}

dim_status solve_batched(dim_csr5* A, const double* b, double* x, const dim_cg_opts& o, dim_cg_result* res,
                         hipStream_t stream) {
    const uint32_t S = A->n_systems;
    const bool transpose = (o.flags & DIM_CG_TRANSPOSE) != 0;
    // block tables: one or more blocks per system, none shared
    std::vector<uint32_t> blk_sys, blk_r0, blk_r1, sys_blk(1, 0);
    for (uint32_t s = 0; s < S; ++s) {
        const uint32_t a = A->sys_row[s], e = A->sys_row[s + 1];
        const uint32_t nb = std::max<uint32_t>(1, (e - a + kRowsPerBlock - 1) / kRowsPerBlock);
        for (uint32_t i = 0; i < nb; ++i) {
            blk_sys.push_back(s);
            blk_r0.push_back(std::min(e, a + i * kRowsPerBlock));
            blk_r1.push_back(std::min(e, a + (i + 1) * kRowsPerBlock));
        }
        sys_blk.push_back(static_cast<uint32_t>(blk_sys.size()));
    }
    const uint32_t B = static_cast<uint32_t>(blk_sys.size());
    const size_t n = A->f.rows;
    const uint32_t check = o.check_every ? std::max<uint32_t>(2, (o.check_every + 1) & ~1u) : 16;

    DevMem m;
    Tables tb{};
    uint32_t *d_bs, *d_r0, *d_r1, *d_sb;
    DIM_HIP_TRY(m.upload(d_bs, blk_sys));
    DIM_HIP_TRY(m.upload(d_r0, blk_r0));
    DIM_HIP_TRY(m.upload(d_r1, blk_r1));
    DIM_HIP_TRY(m.upload(d_sb, sys_blk));
    tb.blk_sys = d_bs;
    tb.blk_r0 = d_r0;
    tb.blk_r1 = d_r1;
    tb.sys_blk = d_sb;
    // the system of each tile, for the product's skip of stopped systems
    std::vector<uint32_t> tile_sys(A->f.tiles);
    for (uint32_t s = 0; s < S; ++s)
        for (uint32_t t = A->sys_tile[s]; t < A->sys_tile[s + 1]; ++t)
            tile_sys[t] = s;
    uint32_t* d_ts = nullptr;
    DIM_HIP_TRY(m.upload(d_ts, tile_sys));
    BState* st;
    SysArrays sa{};
    double *q, *r, *p, *rr0, *rr1, *pq, *bb, *rt;
    DIM_HIP_TRY(m.alloc(st, 1));
    DIM_HIP_TRY(m.alloc(sa.sys_done, S));
    DIM_HIP_TRY(m.alloc(sa.sys_iters, S));
    DIM_HIP_TRY(m.alloc(sa.tol2, S));
    DIM_HIP_TRY(m.alloc(sa.bnorm, S));
    DIM_HIP_TRY(m.alloc(sa.rr, S));
    DIM_HIP_TRY(m.alloc(sa.alpha, S));
    DIM_HIP_TRY(m.alloc(q, n));
    DIM_HIP_TRY(m.alloc(r, n));
    DIM_HIP_TRY(m.alloc(p, n));
    DIM_HIP_TRY(m.alloc(rr0, B));
    DIM_HIP_TRY(m.alloc(rr1, B));
    DIM_HIP_TRY(m.alloc(pq, B));
    DIM_HIP_TRY(m.alloc(bb, B));
    DIM_HIP_TRY(m.alloc(rt, B));

    // capture a batch of iterations (even count: r.r partials alternate)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    DIM_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    auto enqueue = [&](hipStream_t s) {
        for (uint32_t it = 0; it < check; ++it) {
            double* rin = (it & 1) ? rr1 : rr0;
            double* rout = (it & 1) ? rr0 : rr1;
            (void)product(A, transpose, p, q, &st->done, s, d_ts, sa.sys_done);
            hipLaunchKernelGGL(k_b_dot, dim3(B), dim3(kThreads), 0, s, st, sa, tb, static_cast<const double*>(p),
                               static_cast<const double*>(q), pq);
            hipLaunchKernelGGL(k_b_update_r, dim3(B), dim3(kThreads), 0, s, st, sa, tb, r,
                               static_cast<const double*>(q), static_cast<const double*>(pq),
                               static_cast<const double*>(rin), rout);
            hipLaunchKernelGGL(k_b_update_xp, dim3(B), dim3(kThreads), 0, s, st, sa, tb, x,
                               static_cast<const double*>(r), p, static_cast<const double*>(rout),
                               static_cast<const double*>(rin));
        }
    };
    bool use_graph = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (use_graph) {
        enqueue(cap);
        use_graph = hipStreamEndCapture(cap, &graph) == hipSuccess && graph &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        (void)hipGetLastError();
    }
    (void)hipStreamDestroy(cap);
    struct GraphGuard {
        hipGraph_t g;
        hipGraphExec_t e;
        ~GraphGuard() {
            if (e)
                (void)hipGraphExecDestroy(e);
            if (g)
                (void)hipGraphDestroy(g);
        }
    } guard{graph, exec};

    BState* hs = nullptr;
    DIM_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&hs), sizeof(BState), hipHostMallocDefault));
    std::unique_ptr<BState, void (*)(BState*)> hs_guard(hs, [](BState* q) { (void)hipHostFree(q); });

    hipEvent_t e0, e1;
    DIM_HIP_TRY(hipEventCreate(&e0));
    DIM_HIP_TRY(hipEventCreate(&e1));
    DIM_HIP_TRY(hipEventRecord(e0, stream));
    BState init{0, 0, static_cast<int>(o.max_iters), static_cast<int>(S)};
    DIM_HIP_TRY(hipMemcpyAsync(st, &init, sizeof(BState), hipMemcpyHostToDevice, stream));
    DIM_HIP_TRY(product(A, transpose, x, q, nullptr, stream));
    hipLaunchKernelGGL(k_b_init, dim3(B), dim3(kThreads), 0, stream, tb, b, static_cast<const double*>(q), r, p, rr0, bb);
    hipLaunchKernelGGL(k_b_init_state, dim3(1), dim3(64), 0, stream, st, sa, tb, static_cast<const double*>(rr0),
                       static_cast<const double*>(bb), o.rtol);
    DIM_HIP_TRY(hipGetLastError());
    for (;;) {
        DIM_HIP_TRY(hipMemcpyAsync(hs, st, sizeof(BState), hipMemcpyDeviceToHost, stream));
        DIM_HIP_TRY(hipStreamSynchronize(stream));
        if (hs->done)
            break;
        if (use_graph) {
            DIM_HIP_TRY(hipGraphLaunch(exec, stream));
        } else {
            enqueue(stream);
            DIM_HIP_TRY(hipGetLastError());
        }
    }
    DIM_HIP_TRY(hipEventRecord(e1, stream));
    DIM_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    // FP64 true residual of every system
    DIM_HIP_TRY(product(A, transpose, x, q, nullptr, stream));
    hipLaunchKernelGGL(k_b_resid, dim3(B), dim3(kThreads), 0, stream, tb, b, static_cast<const double*>(q), rt);
    DIM_HIP_TRY(hipGetLastError());
    std::vector<double> parts(B), h_rr(S), h_bn(S);
    std::vector<int> h_done(S), h_it(S);
    DIM_HIP_TRY(hipMemcpyAsync(parts.data(), rt, B * sizeof(double), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipMemcpyAsync(h_rr.data(), sa.rr, S * sizeof(double), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipMemcpyAsync(h_bn.data(), sa.bnorm, S * sizeof(double), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipMemcpyAsync(h_done.data(), sa.sys_done, S * sizeof(int), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipMemcpyAsync(h_it.data(), sa.sys_iters, S * sizeof(int), hipMemcpyDeviceToHost, stream));
    DIM_HIP_TRY(hipStreamSynchronize(stream));

    dim_status status = DIM_SUCCESS;
    for (uint32_t s = 0; s < S; ++s) {
        double tr = 0.0;
        for (uint32_t i = sys_blk[s]; i < sys_blk[s + 1]; ++i)
            tr += parts[i];
        const double scale = h_bn[s] > 0.0 ? h_bn[s] : 1.0;
        dim_cg_result& R = res[s];
        R.converged = h_done[s] == 1;
        R.iterations = static_cast<uint32_t>(h_done[s] ? h_it[s] : hs->iterations);
        R.relres = std::sqrt(h_rr[s]) / scale;
        R.relres_true = std::sqrt(tr) / scale;
        R.refinements = 0;
        R.seconds = ms * 1e-3;
        if (h_done[s] == 2)
            status = DIM_BREAKDOWN;
        else if (!R.converged && status == DIM_SUCCESS)
            status = DIM_NOT_CONVERGED;
    }
    if (status == DIM_BREAKDOWN)
        set_error("CG breakdown in at least one system: p'Ap <= 0 or not finite");
    else if (status == DIM_NOT_CONVERGED)
        set_error("CG reached max_iters above the tolerance in at least one system");
    return status;
}

// Times the candidate tile heights on the whole batch and keeps the fastest.
dim_status autotune_batched(const std::vector<SystemCsr>& sys, bool keep_f32, dim_csr5** out) {
    uint64_t rows = 0, nnz = 0;
    for (const SystemCsr& s : sys) {
        rows += s.rows;
        nnz += s.nnz;
    }
    dim_csr5* best = nullptr;
    float best_ms = 0;
    double *x = nullptr, *y = nullptr;
    DIM_HIP_TRY(hipMalloc(&x, std::max<uint64_t>(rows, 1) * sizeof(double)));
    DIM_HIP_TRY(hipMalloc(&y, std::max<uint64_t>(rows, 1) * sizeof(double)));
    std::unique_ptr<double, void (*)(double*)> gx(x, [](double* q) { (void)hipFree(q); });
    std::unique_ptr<double, void (*)(double*)> gy(y, [](double* q) { (void)hipFree(q); });
    {
        std::vector<double> hx(std::max<uint64_t>(rows, 1));
        std::mt19937_64 rng(7);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (auto& v : hx)
            v = U(rng);
        DIM_HIP_TRY(hipMemcpy(x, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    hipEvent_t e0, e1;
    DIM_HIP_TRY(hipEventCreate(&e0));
    DIM_HIP_TRY(hipEventCreate(&e1));
    dim_status st = DIM_SUCCESS;
    for (uint32_t sigma : sigma_candidates(static_cast<uint32_t>(std::min<uint64_t>(rows, 0xffffffffu)), nnz)) {
        HostFormat f;
        std::vector<uint32_t> sr, stl;
        std::vector<uint64_t> sn;
        std::string err = build_batched(sys, sigma, f, sr, stl, sn);
        if (!err.empty()) {
            st = fail(DIM_INVALID_ARGUMENT, err);
            break;
        }
        dim_csr5* h = nullptr;
        if ((st = finish_batched(f, keep_f32, sr, stl, sn, &h)) != DIM_SUCCESS)
            break;
        for (int i = 0; i < 2; ++i)
            (void)launch_spmv(h->f, h->s, false, 1, x, y, 0, h->f.tiles, 0, 0, h->f.n_empty, false, nullptr, 0);
        (void)hipEventRecord(e0, 0);
        for (int i = 0; i < 8; ++i)
            (void)launch_spmv(h->f, h->s, false, 1, x, y, 0, h->f.tiles, 0, 0, h->f.n_empty, false, nullptr, 0);
        (void)hipEventRecord(e1, 0);
        hipError_t e = hipEventSynchronize(e1);
        if (e != hipSuccess) {
            free_handle(h);
            st = hip_fail(e, "autotune timing");
            break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (!best || ms < best_ms) {
            free_handle(best);
            best = h;
            best_ms = ms;
        } else {
            free_handle(h);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != DIM_SUCCESS) {
        free_handle(best);
        return st;
    }
    *out = best;
    return DIM_SUCCESS;
}

} // namespace
} // namespace dimgpu

using namespace dimgpu;

extern "C" {

dim_status dim_csr5_from_csr_batched(uint32_t n_systems, const uint32_t* rows, const uint64_t* nnz,
                                     const uint32_t* const* row_ptr, const uint32_t* const* col_idx,
                                     const double* const* vals, const dim_csr5_opts* opts, dim_csr5** out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    if (n_systems == 0 || !rows || !nnz || !row_ptr || !col_idx || !vals)
        return fail(DIM_INVALID_ARGUMENT, "n_systems must be positive and every array given");
    std::vector<SystemCsr> sys(n_systems);
    for (uint32_t i = 0; i < n_systems; ++i) {
        sys[i] = SystemCsr{rows[i], nnz[i], row_ptr[i], col_idx[i], vals[i]};
        if (nnz[i] && !vals[i])
            return fail(DIM_INVALID_ARGUMENT, "vals is null");
        std::string err = validate_csr(rows[i], rows[i], nnz[i], row_ptr[i], col_idx[i]);
        if (!err.empty())
            return fail(DIM_INVALID_ARGUMENT, "system " + std::to_string(i) + ": " + err);
    }
    const uint32_t sigma = opts ? opts->sigma : 16;
    const bool keep_f32 = opts && (opts->flags & DIM_CSR5_KEEP_F32);
    if (sigma == 0)
        return autotune_batched(sys, keep_f32, out);
    HostFormat f;
    std::vector<uint32_t> sr, stl;
    std::vector<uint64_t> sn;
    std::string err = build_batched(sys, sigma, f, sr, stl, sn);
    if (!err.empty())
        return fail(DIM_INVALID_ARGUMENT, err);
    return finish_batched(f, keep_f32, sr, stl, sn, out);
}

dim_status dim_cg_batched(const dim_csr5* A, const double* B, double* X, const dim_cg_opts* opts, dim_cg_result* res,
                          void* stream) {
    if (!A)
        return fail(DIM_INVALID_ARGUMENT, "matrix handle is null");
    if (!opts || !res)
        return fail(DIM_INVALID_ARGUMENT, "opts and res are required");
    if (A->f.rows && (!B || !X))
        return fail(DIM_INVALID_ARGUMENT, "null vector");
    if (opts->precision != DIM_CG_F64)
        return fail(DIM_INVALID_ARGUMENT, "batched CG runs in FP64");
    if (!(opts->rtol >= 0.0))
        return fail(DIM_INVALID_ARGUMENT, "rtol must be >= 0");
    return solve_batched(const_cast<dim_csr5*>(A), B, X, *opts, res, static_cast<hipStream_t>(stream));
}

} // extern "C"
