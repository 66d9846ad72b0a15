// Many independent SPD systems solved together: one block-diagonal CSR5
// matrix (every system padded to whole tiles, so no tile straddles two
// systems) and a CG whose scalars are kept per system.
//
// The matrix carries a chunk table in which no chunk holds tiles of two
// systems, and, when every system qualifies, column codes with one table of
// diagonals per system. One iteration is three kernels on the whole batch:
//   1. q = A p over the chunk table, with p.q of each chunk's rows left as
//      one partial per chunk (chunks of stopped systems are skipped), plus
//      the carry fix-up when some chunk boundary needs it;
//   2. per system: alpha = r.r / p.q (its chunks' partials in a fixed order),
//      r -= alpha q, partials of the new r.r;
//   3. per system: the stop test, x += alpha p, p = r + beta p.
// A batch of iterations is captured once in a hipGraph and replayed until
// every system has stopped; the host looks at the state between batches.
//
// With DIM_CG_PERSISTENT the same three phases run inside one persistent
// cooperative kernel that keeps the whole iteration loop on the device,
// separated by grid barriers (see k_b_persistent); its iterates are bitwise
// those of the launch sequence. Measured on the MI355X with 32 systems
// (5.6M rows, 4792 iterations) it takes 0.81 s against 0.62 s: about 13 us
// more per grid barrier, three per iteration, than a kernel boundary inside
// the graph, so the launch sequence stays the default.
//
// The adjoint solve with A^T products (DIM_CG_TRANSPOSE) uses the transposed
// product and a separate p.q pass. Partial sums of a system are reduced in a
// fixed order, so iterates are reproducible (apart from the transposed
// product, which scatters with atomics).
#include "handle.h"
#include "spmv_impl.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

namespace dimgpu {

namespace {

using detail::sload;
using detail::TileParams;

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

// The block-diagonal format. Column codes are kept when every system has
// them (stencil systems do): the code words are concatenated and col_dict
// then holds kMaxDiagonals diagonals per system (the system's own table; a
// diagonal col - row does not change when rows and columns shift together).
std::string build_batched(const std::vector<SystemCsr>& sys, uint32_t sigma, HostFormat& out,
                          std::vector<uint32_t>& sys_row, std::vector<uint32_t>& sys_tile,
                          std::vector<uint64_t>& sys_nnz) {
    out = HostFormat{};
    out.sigma = sigma;
    sys_row.assign(1, 0);
    sys_tile.assign(1, 0);
    sys_nnz.assign(1, 0);
    uint64_t rows = 0, tiles = 0, tdo_total = 0;
    bool codes = true;
    std::vector<int32_t> dicts;
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
        codes = codes && !f.col_code.empty();
        if (codes) {
            out.col_code.insert(out.col_code.end(), f.col_code.begin(), f.col_code.end());
            std::vector<int32_t> d(f.col_dict);
            d.resize(kMaxDiagonals, 0);
            dicts.insert(dicts.end(), d.begin(), d.end());
        }
        rows += s.rows;
        tiles += f.tiles;
        if (rows >= kDirty || tiles >= kDirty)
            return "batch too large";
        sys_row.push_back(static_cast<uint32_t>(rows));
        sys_tile.push_back(static_cast<uint32_t>(tiles));
        sys_nnz.push_back(sys_nnz.back() + s.nnz);
    }
    if (codes && !sys.empty() && uint64_t{rows} * sizeof(double) <= (uint64_t{1} << 32)) {
        out.col_dict = std::move(dicts);
    } else {
        out.col_code.clear();
        out.col_dict.clear();
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
    double* rr;      // r.r of the last step (kept by k_b_update_xp's block 0)
    double* alpha;   // this iteration's step (0 for a stopped system)
    double* rr_old;  // r.r this iteration started from, for beta
    // persistent solver: the other parity's statuses and r.r, and this
    // iteration's breakdown flags
    int* done_alt;
    double* rr_alt;
    int* broke;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// sum of parts[b0..b1) by wave 0, fixed order (lane l adds entries l, l + 64,
// ..., then the butterfly); result broadcast through *out
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

// sum of parts[b0..b1) by the whole block (long ranges: a system's per-chunk
// p.q partials): thread t adds entries t, t + blockDim.x, ..., each wave
// reduces by butterfly and thread 0 adds the waves in order. The order
// depends only on the range and the block size, so every block of a system
// obtains the same bits. Result broadcast through *out.
__device__ void block_range_sum(const double* parts, uint32_t b0, uint32_t b1, double* out) {
    __shared__ double red[kThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double s = 0.0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
        s += parts[i];
    s = wave_sum(s);
    if (lane == 0)
        red[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < static_cast<int>(blockDim.x >> 6); ++i)
            t += red[i];
        *out = t;
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
// pq_part holds the p.q partials, those of system s at [pq_range[s],
// pq_range[s + 1]) (one per chunk from the fused product, or one per block
// from k_b_dot). r.r is sa.rr, the sum the previous k_b_update_xp (or
// k_b_init_state) kept; the system's first block copies it to sa.rr_old for
// beta, since k_b_update_xp's block 0 overwrites sa.rr.
__global__ __launch_bounds__(kThreads) void k_b_update_r(BState* st, SysArrays sa, Tables tb, double* r,
                                                         const double* q, const double* pq_part,
                                                         const uint32_t* pq_range, double* rr_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        st->done_prev = st->done;
    if (st->done)
        return;
    const uint32_t s = tb.blk_sys[blockIdx.x];
    __shared__ double pq, alpha;
    __shared__ int skip;
    block_range_sum(pq_part, pq_range[s], pq_range[s + 1], &pq);
    if (threadIdx.x == 0) {
        const double rr = sa.rr[s];
        if (blockIdx.x == tb.sys_blk[s])
            sa.rr_old[s] = rr;
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
                                                          const double* r, double* p, const double* rr_new_part) {
    if (st->done_prev)
        return;
    const uint32_t s = tb.blk_sys[blockIdx.x];
    __shared__ double rn;
    range_sum(rr_new_part, tb.sys_blk[s], tb.sys_blk[s + 1], &rn);
    const double ro = sa.rr_old[s];
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

// ---- the persistent solver (DIM_CG_PERSISTENT) ---------------------------

constexpr int kPThreads = 1024;          // threads per block of the persistent kernel (16 wavefronts)
constexpr int kPWaves = kPThreads / 64;
// block_range_sum's order depends on the block size: the persistent kernel
// sums p.q like k_b_update_r only with blocks of the same size (its work
// items are the launch sequence's blocks of kRowsPerBlock rows)
static_assert(kPThreads == kThreads, "persistent and launch-sequence blocks differ");
constexpr uint32_t kFlagStride = 32;     // barrier flags 128 bytes apart

struct PersistParams {
    TileParams tp;              // q = A p with p.q per chunk over the chunk table
    Tables tb;                  // vector work items: the launch sequence's blocks
    SysArrays sa;
    const uint32_t* sys_chunk;  // chunks of system s: [sys_chunk[s], sys_chunk[s + 1])
    double *x, *r, *p, *q;
    double* rr_part;            // one partial per work item
    double* dot_chunk;          // one partial per chunk
    int fixup;                  // some chunk boundary needs the carry fix-up
    uint32_t nchunks, n_items, n_sys;
    int max_iters;
    BState* st;
    // barrier words (zero before the launch): [0] the generation block 0
    // releases, [kFlagStride * (1 + b)] the generation block b has reached
    unsigned* bar;
    int* abort;                 // set when a barrier waits beyond bar_limit
    uint64_t bar_limit;         // wall-clock ticks
};

// Waits until *w == target, at most `limit` wall-clock ticks after t0 or
// until another block has set the abort flag; false when the wait failed (and
// the abort flag is then set).
__device__ bool wait_word(const unsigned* w, unsigned target, int* abort, uint64_t t0, uint64_t limit) {
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
        if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            static_cast<uint64_t>(wall_clock64()) - t0 > limit) {
            atomicExch(abort, 1);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// Grid barrier of the cooperative launch (every block resident), built so
// that no word sees more than one writer per generation: thread 0 of each
// block publishes the block's writes at device scope (release fence) and
// stores the generation in the block's own flag line; wavefront 0 of block 0
// waits for every flag, then releases the generation word; the other blocks
// wait for that word and acquire. Every wait is bounded in time. Returns
// false when the solve must stop; the result is uniform across the block.
__device__ bool grid_barrier(const PersistParams& P, unsigned& gen) {
    __shared__ int ok_s;
    __syncthreads();
    const unsigned target = gen + 1;
    gen = target;
    int ok = 1;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(P.bar + kFlagStride * (1u + blockIdx.x), target, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x < 64) {
            const uint64_t t0 = wall_clock64();
            for (uint32_t b = threadIdx.x; b < gridDim.x && ok; b += 64)
                ok = wait_word(P.bar + kFlagStride * (1u + b), target, P.abort, t0, P.bar_limit);
            ok = __ballot(ok) == ~0ull;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            if (threadIdx.x == 0) {
                if (ok)
                    __hip_atomic_store(P.bar, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok_s = ok;
            }
        }
    } else if (threadIdx.x == 0) {
        ok = wait_word(P.bar, target, P.abort, wall_clock64(), P.bar_limit);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok_s = ok;
    }
    __syncthreads();
    return ok_s != 0;
}

// wave 0 sums parts[b0..b1) in a fixed order (lane l: l, l + 64, ..., then
// the butterfly): every block obtains the same bits. Result in *out (LDS).
__device__ void range_sum_p(const double* parts, uint32_t b0, uint32_t b1, double* out) {
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

__device__ void block_sum_store_p(double acc, double* dst) {
    __shared__ double red[kPWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    acc = wave_sum(acc);
    if (lane == 0)
        red[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kPWaves; ++i)
            t += red[i];
        *dst = t;
    }
    __syncthreads();
}

// The whole iteration loop of the batched CG in one launch, three phases per
// iteration separated by grid barriers:
//   A. q = A p over the chunk table, p.q per chunk (chunks of stopped systems
//      skipped), then the carry fix-up when a chunk boundary needs it;
//   B. per work item of a running system: alpha = r.r / p.q, r -= alpha q,
//      the item's r.r partial;
//   C. per work item: the stop test, x += alpha p, p = r + beta p.
// Statuses and r.r are double-buffered by iteration parity: iteration i
// reads `cur` and the first item of each system writes `nxt`, so no phase
// reads what the same phase writes.
template<int SIGMA, bool CODES>
__global__ __launch_bounds__(kPThreads) void k_b_persistent(PersistParams P) {
    const int lane = threadIdx.x & 63;
    const uint32_t nwaves = gridDim.x * kPWaves;
    const uint32_t wid = detail::xcd_remap(blockIdx.x, gridDim.x) * kPWaves + (threadIdx.x >> 6);
    __shared__ double sh_a, sh_b;
    __shared__ int all_s;
    unsigned gen = 0;
    int it = 0;
    int* cur = P.sa.sys_done;
    int* nxt = P.sa.done_alt;
    double* rr_cur = P.sa.rr;
    double* rr_nxt = P.sa.rr_alt;
    int all = 0;
    for (;;) {
        // ---- A: q = A p over the chunks, p.q per chunk
        for (uint32_t c0 = wid; c0 < P.nchunks; c0 += nwaves) {
            const uint32_t c = __builtin_amdgcn_readfirstlane(c0);
            const uint32_t s = sload(P.tp.chunk_sys + c);
            const bool stopped = __builtin_amdgcn_readfirstlane(cur[s]) != 0;
            const double d = detail::csr5_chunk<double, SIGMA, 1, true, CODES>(P.tp, c, stopped);
            if (lane == 0)
                P.dot_chunk[c] = d;
        }
        if (P.fixup) {
            // rows that cross chunk boundaries the look-ahead cannot finish:
            // the chunk in which such a row ends adds the chain in chunk order
            if (!grid_barrier(P, gen))
                return;
            const double* head = static_cast<const double*>(P.tp.head);
            const double* tail = static_cast<const double*>(P.tp.tail);
            for (uint32_t c = blockIdx.x * kPThreads + threadIdx.x; c < P.nchunks; c += gridDim.x * kPThreads) {
                if (!(P.tp.chunk_flags[c] & detail::kChunkHead))
                    continue;
                int64_t kk = static_cast<int64_t>(c) - 1;
                while (kk >= 0 && !(P.tp.chunk_flags[kk] & detail::kChunkRowStart))
                    --kk;
                const uint32_t from = kk >= 0 ? static_cast<uint32_t>(kk) : 0u;
                const uint32_t row = P.tp.tile_ptr[P.tp.chunk_tile[c]] & ~kDirty;
                double v;
                if (from < c) {
                    v = tail[from];
                    for (uint32_t u = from + 1; u < c; ++u)
                        v = v + tail[u];
                    v = v + head[c];
                } else {
                    v = head[c];
                }
                P.q[row] = v;
                P.dot_chunk[c] += P.p[row] * v;
            }
        }
        if (!grid_barrier(P, gen))
            return;
        // ---- B: alpha, r -= alpha q, r.r partials
        for (uint32_t item = blockIdx.x; item < P.n_items; item += gridDim.x) {
            const uint32_t s = P.tb.blk_sys[item];
            if (cur[s]) { // block-uniform: cur does not change in this iteration
                if (threadIdx.x == 0)
                    P.rr_part[item] = 0.0;
                continue;
            }
            block_range_sum(P.dot_chunk, P.sys_chunk[s], P.sys_chunk[s + 1], &sh_a);
            if (threadIdx.x == 0) {
                const double pq = sh_a;
                double a = rr_cur[s] / pq;
                const bool broke = !(pq > 0.0) || !std::isfinite(a);
                if (broke)
                    a = 0.0;
                sh_b = a;
                if (item == P.tb.sys_blk[s]) {
                    P.sa.alpha[s] = a;
                    P.sa.broke[s] = broke ? 1 : 0;
                }
            }
            __syncthreads();
            const double a = sh_b;
            double acc = 0.0;
            const uint32_t r1 = P.tb.blk_r1[item];
            for (uint32_t i = P.tb.blk_r0[item] + threadIdx.x; i < r1; i += kPThreads) {
                const double ri = P.r[i] - a * P.q[i];
                P.r[i] = ri;
                acc += ri * ri;
            }
            block_sum_store_p(acc, P.rr_part + item);
        }
        if (!grid_barrier(P, gen))
            return;
        // ---- C: stop test, x += alpha p, p = r + beta p, next statuses
        for (uint32_t item = blockIdx.x; item < P.n_items; item += gridDim.x) {
            const uint32_t s = P.tb.blk_sys[item];
            const bool first = item == P.tb.sys_blk[s];
            const int status = cur[s];
            if (status) {
                if (first && threadIdx.x == 0) {
                    nxt[s] = status;
                    rr_nxt[s] = rr_cur[s];
                }
                continue;
            }
            range_sum_p(P.rr_part, P.tb.sys_blk[s], P.tb.sys_blk[s + 1], &sh_a);
            const double rn = sh_a;
            const double a = P.sa.alpha[s];
            const bool broke = P.sa.broke[s] != 0;
            const bool conv = rn <= P.sa.tol2[s];
            const double beta = rn / rr_cur[s];
            if (!broke) {
                const uint32_t r1 = P.tb.blk_r1[item];
                for (uint32_t i = P.tb.blk_r0[item] + threadIdx.x; i < r1; i += kPThreads) {
                    const double pi = P.p[i];
                    P.x[i] = P.x[i] + a * pi;
                    if (!conv)
                        P.p[i] = P.r[i] + beta * pi;
                }
            }
            if (first && threadIdx.x == 0) {
                if (broke) {
                    nxt[s] = 2;
                    rr_nxt[s] = rr_cur[s];
                    P.sa.sys_iters[s] = it;
                } else {
                    nxt[s] = conv ? 1 : 0;
                    rr_nxt[s] = rn;
                    if (conv)
                        P.sa.sys_iters[s] = it + 1;
                }
            }
            __syncthreads(); // sh_a is reused by the next item
        }
        ++it;
        if (!grid_barrier(P, gen))
            return;
        int* t = cur;
        cur = nxt;
        nxt = t;
        double* u = rr_cur;
        rr_cur = rr_nxt;
        rr_nxt = u;
        // every block reads the same statuses and decides alike
        if (threadIdx.x == 0)
            all_s = 1;
        __syncthreads();
        for (uint32_t s = threadIdx.x; s < P.n_sys; s += kPThreads)
            if (!cur[s])
                atomicAnd(&all_s, 0);
        __syncthreads();
        all = all_s;
        if (all || it >= P.max_iters)
            break;
    }
    // block 0 leaves the final statuses and r.r in sys_done / rr (the other
    // blocks only read cur, which is not overwritten here)
    if (blockIdx.x == 0) {
        for (uint32_t s = threadIdx.x; s < P.n_sys; s += kPThreads) {
            P.sa.sys_done[s] = cur[s];
            P.sa.rr[s] = rr_cur[s];
        }
        if (threadIdx.x == 0) {
            P.st->iterations = it;
            P.st->done = all ? 1 : 3;
        }
    }
}

template<int SIGMA, bool CODES>
int persistent_blocks_per_cu() {
    static int b = -1;
    if (b < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_b_persistent<SIGMA, CODES>, kPThreads, 0) !=
              hipSuccess ||
            n < 0)
            n = 0;
        b = n;
    }
    return b;
}

// Resident blocks per CU of the persistent kernel for this format (0: no
// kernel for it).
int persistent_occupancy(uint32_t sigma, bool codes) {
    switch (sigma) {
#define DIM_OCC_CASE(S)                                                                                               \
    case S:                                                                                                           \
        if constexpr (S <= static_cast<int>(kMaxCodedSigma))                                                          \
            if (codes)                                                                                                \
                return persistent_blocks_per_cu<S, true>();                                                           \
        return persistent_blocks_per_cu<S, false>();
        DIM_FOR_EACH_SIGMA(DIM_OCC_CASE)
#undef DIM_OCC_CASE
        default: return 0;
    }
}

hipError_t launch_persistent(uint32_t sigma, bool codes, uint32_t grid, PersistParams& P, hipStream_t s) {
    void* args[] = {&P};
    switch (sigma) {
#define DIM_PLAUNCH_CASE(S)                                                                                           \
    case S:                                                                                                           \
        if constexpr (S <= static_cast<int>(kMaxCodedSigma))                                                          \
            if (codes)                                                                                                \
                return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_b_persistent<S, true>), dim3(grid), \
                                                  dim3(kPThreads), args, 0, s);                                       \
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_b_persistent<S, false>), dim3(grid),      \
                                          dim3(kPThreads), args, 0, s);
        DIM_FOR_EACH_SIGMA(DIM_PLAUNCH_CASE)
#undef DIM_PLAUNCH_CASE
        default: return hipErrorInvalidValue;
    }
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

// The chunk table of a batch: about one chunk per resident wavefront of the
// product with its dot epilogue, never two systems in one chunk. Sizes are balanced
// within each system; at most `waves` chunks in all when waves exceeds the
// number of systems.
void build_chunk_table(const std::vector<uint32_t>& sys_tile, uint64_t waves, std::vector<uint32_t>& chunk_tile,
                       std::vector<uint32_t>& chunk_sys, std::vector<uint32_t>& sys_chunk) {
    const uint32_t S = static_cast<uint32_t>(sys_tile.size() - 1);
    const uint64_t T = sys_tile.back();
    const uint64_t usable = waves > S ? waves - S : 1;
    const uint64_t tpw = std::max<uint64_t>(1, (T + usable - 1) / usable);
    chunk_tile.clear();
    chunk_sys.clear();
    sys_chunk.assign(1, 0);
    for (uint32_t s = 0; s < S; ++s) {
        const uint64_t a = sys_tile[s], n = sys_tile[s + 1] - a;
        const uint64_t nc = std::max<uint64_t>(1, (n + tpw - 1) / tpw);
        const uint64_t base = n / nc, rem = n % nc;
        uint64_t t = a;
        for (uint64_t i = 0; i < nc; ++i) {
            chunk_tile.push_back(static_cast<uint32_t>(t));
            chunk_sys.push_back(s);
            t += base + (i < rem ? 1 : 0);
        }
        sys_chunk.push_back(static_cast<uint32_t>(chunk_sys.size()));
    }
    chunk_tile.push_back(static_cast<uint32_t>(T));
}

dim_status finish_batched(HostFormat& f, bool keep_f32, const std::vector<uint32_t>& sys_row,
                          const std::vector<uint32_t>& sys_tile, const std::vector<uint64_t>& sys_nnz, dim_csr5** out) {
    // upload_format keeps one dictionary; a batch keeps one per system
    std::vector<int32_t> dicts;
    dicts.swap(f.col_dict);
    const bool codes = !dicts.empty() && !f.col_code.empty();
    if (codes)
        f.col_dict.assign(dicts.begin(), dicts.begin() + kMaxDiagonals);
    dim_csr5* h = nullptr;
    if (dim_status st = upload_format(f, keep_f32, &h))
        return st;
    std::unique_ptr<dim_csr5, void (*)(dim_csr5*)> g(h, free_handle);
    h->n_systems = static_cast<uint32_t>(sys_row.size() - 1);
    h->sys_row = sys_row;
    h->sys_tile = sys_tile;
    h->sys_nnz_csr = sys_nnz;
    h->nnz_logical = sys_nnz.back();
    if (codes) {
        int32_t* d = nullptr;
        DIM_HIP_TRY(upload(h, d, dicts));
        h->f.col_dict = d;
    }
    // chunk table: about one chunk per resident wavefront of the product
    // with its dot epilogue
    const uint64_t waves = detail::resident_tile_waves(f.sigma, h->f.col_code != nullptr, true);
    std::vector<uint32_t> ct, cs;
    build_chunk_table(sys_tile, waves, ct, cs, h->sys_chunk);
    uint32_t *dct = nullptr, *dcs = nullptr, *dsc = nullptr;
    DIM_HIP_TRY(upload(h, dct, ct));
    DIM_HIP_TRY(upload(h, dcs, cs));
    DIM_HIP_TRY(upload(h, dsc, h->sys_chunk));
    h->f.chunk_tile = dct;
    h->f.chunk_sys = dcs;
    h->d_sys_chunk = dsc;
    h->f.n_chunk_tab = static_cast<uint32_t>(cs.size());
    h->f.chunk_tab_fixup = h->binfo.needs_fixup_table(ct);
    *out = g.release();
    return DIM_SUCCESS;
}

// sys_done (may be null): chunks of stopped systems are skipped.
hipError_t product(dim_csr5* A, bool transpose, const double* x, double* y, const int* done, hipStream_t s,
                   const int* sys_done = nullptr) {
    if (transpose)
        return launch_spmv_transpose(A->f, false, x, y, s, done);
    return launch_spmv(A->f, A->s, false, 1, x, y, 0, A->f.tiles, 0, 0, A->f.n_empty, false, done, s, nullptr, 0, 0,
                       nullptr, sys_done);
}

// Vector work items: blocks of at most `rows_per` rows of one system.
void make_items(const dim_csr5* A, uint32_t rows_per, std::vector<uint32_t>& blk_sys, std::vector<uint32_t>& blk_r0,
                std::vector<uint32_t>& blk_r1, std::vector<uint32_t>& sys_blk) {
    blk_sys.clear();
    blk_r0.clear();
    blk_r1.clear();
    sys_blk.assign(1, 0);
    for (uint32_t s = 0; s < A->n_systems; ++s) {
        const uint32_t a = A->sys_row[s], e = A->sys_row[s + 1];
        const uint32_t nb = std::max<uint32_t>(1, (e - a + rows_per - 1) / rows_per);
        for (uint32_t i = 0; i < nb; ++i) {
            blk_sys.push_back(s);
            blk_r0.push_back(std::min(e, a + i * rows_per));
            blk_r1.push_back(std::min(e, a + (i + 1) * rows_per));
        }
        sys_blk.push_back(static_cast<uint32_t>(blk_sys.size()));
    }
}

dim_status report(dim_csr5* A, const double* b, double* x, bool transpose, const Tables& tb,
                  const std::vector<uint32_t>& sys_blk, uint32_t B, double* q, double* rt, const SysArrays& sa,
                  uint32_t threads, const BState& hs, float ms, dim_cg_result* res, hipStream_t stream) {
    const uint32_t S = A->n_systems;
    // FP64 true residual of every system
    DIM_HIP_TRY(product(A, transpose, x, q, nullptr, stream));
    (void)threads;
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
        R.iterations = static_cast<uint32_t>(h_done[s] ? h_it[s] : hs.iterations);
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

} // namespace

// The buffers of a batched solve, all sized by the batch alone: the block
// tables, the state, the per-system arrays, q, r, p and the partials, the
// pinned copy of the state, the capture stream and the captured batch of
// iterations. A handle keeps them (batch_cache) from its first solve when
// q, r and p take at most kBatchCacheMaxBytes; the graph is reused while
// the direction (A or A^T), the batch length and x stay the same. Every
// buffer is written before it is read.
struct BatchCache {
    DevMem m;
    std::vector<uint32_t> blk_sys, blk_r0, blk_r1, sys_blk;
    uint32_t B = 0;
    Tables tb{};
    uint32_t* d_sb = nullptr;
    BState* st = nullptr;
    SysArrays sa{};
    double *q = nullptr, *r = nullptr, *p = nullptr, *rr = nullptr, *pq = nullptr, *bb = nullptr, *rt = nullptr;
    BState* hs = nullptr; // pinned
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool g_transpose = false;
    uint32_t g_check = 0;
    const double* g_x = nullptr;
    void reset_graph() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~BatchCache() {
        reset_graph();
        if (cap)
            (void)hipStreamDestroy(cap);
        if (hs)
            (void)hipHostFree(hs);
    }
};

void free_batch_cache(BatchCache* c) {
    delete c;
}

namespace {

constexpr size_t kBatchCacheMaxBytes = size_t{256} << 20;

dim_status init_batch_cache(dim_csr5* A, BatchCache& c) {
    const uint32_t S = A->n_systems;
    const size_t n = A->f.rows;
    // block tables: one or more blocks of kRowsPerBlock rows per system, none
    // shared (the vector kernels, the residuals and the persistent kernel's
    // work items)
    make_items(A, kRowsPerBlock, c.blk_sys, c.blk_r0, c.blk_r1, c.sys_blk);
    c.B = static_cast<uint32_t>(c.blk_sys.size());
    const uint32_t B = c.B;
    DevMem& m = c.m;
    uint32_t *d_bs, *d_r0, *d_r1;
    DIM_HIP_TRY(m.upload(d_bs, c.blk_sys));
    DIM_HIP_TRY(m.upload(d_r0, c.blk_r0));
    DIM_HIP_TRY(m.upload(d_r1, c.blk_r1));
    DIM_HIP_TRY(m.upload(c.d_sb, c.sys_blk));
    c.tb.blk_sys = d_bs;
    c.tb.blk_r0 = d_r0;
    c.tb.blk_r1 = d_r1;
    c.tb.sys_blk = c.d_sb;
    SysArrays& sa = c.sa;
    DIM_HIP_TRY(m.alloc(c.st, 1));
    DIM_HIP_TRY(m.alloc(sa.sys_done, S));
    DIM_HIP_TRY(m.alloc(sa.sys_iters, S));
    DIM_HIP_TRY(m.alloc(sa.tol2, S));
    DIM_HIP_TRY(m.alloc(sa.bnorm, S));
    DIM_HIP_TRY(m.alloc(sa.rr, S));
    DIM_HIP_TRY(m.alloc(sa.alpha, S));
    DIM_HIP_TRY(m.alloc(sa.rr_old, S));
    DIM_HIP_TRY(m.alloc(sa.done_alt, S));
    DIM_HIP_TRY(m.alloc(sa.rr_alt, S));
    DIM_HIP_TRY(m.alloc(sa.broke, S));
    DIM_HIP_TRY(m.alloc(c.q, n));
    DIM_HIP_TRY(m.alloc(c.r, n));
    DIM_HIP_TRY(m.alloc(c.p, n));
    DIM_HIP_TRY(m.alloc(c.rr, B));
    DIM_HIP_TRY(m.alloc(c.pq, B));
    DIM_HIP_TRY(m.alloc(c.bb, B));
    DIM_HIP_TRY(m.alloc(c.rt, B));
    DIM_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&c.hs), sizeof(BState), hipHostMallocDefault));
    DIM_HIP_TRY(hipStreamCreateWithFlags(&c.cap, hipStreamNonBlocking));
    return DIM_SUCCESS;
}

dim_status solve_batched(dim_csr5* A, const double* b, double* x, const dim_cg_opts& o, dim_cg_result* res,
                         hipStream_t stream) {
    const uint32_t S = A->n_systems;
    const bool transpose = (o.flags & DIM_CG_TRANSPOSE) != 0;
    const bool codes = A->f.col_code != nullptr;
    const bool persistent = (o.flags & DIM_CG_PERSISTENT) != 0;
    const int occ = persistent ? persistent_occupancy(A->f.sigma, codes) : 0;
    if (persistent && occ < 1)
        return fail(DIM_UNSUPPORTED, "batched CG: the persistent kernel cannot be resident for this format");
    const size_t n = A->f.rows;
    const uint32_t check = o.check_every ? o.check_every : 16;
    // the handle's buffers (kept from its first solve), or this call's
    BatchCache local;
    BatchCache* bc = &local;
    if (3 * n * sizeof(double) <= kBatchCacheMaxBytes) {
        if (!A->batch_cache) {
            std::unique_ptr<BatchCache> fresh(new BatchCache);
            if (dim_status e = init_batch_cache(A, *fresh))
                return e;
            A->batch_cache = fresh.release();
        }
        bc = A->batch_cache;
    } else if (dim_status e = init_batch_cache(A, local)) {
        return e;
    }
    const uint32_t B = bc->B;
    const Tables tb = bc->tb;
    const std::vector<uint32_t>& sys_blk = bc->sys_blk;
    uint32_t* d_sb = bc->d_sb;
    BState* st = bc->st;
    const SysArrays sa = bc->sa;
    double *q = bc->q, *r = bc->r, *p = bc->p, *rr = bc->rr, *pq = bc->pq, *bb = bc->bb, *rt = bc->rt;
    BState* hs = bc->hs;
    DevMem m; // the persistent kernel's own buffers

    hipEvent_t e0, e1;
    DIM_HIP_TRY(hipEventCreate(&e0));
    DIM_HIP_TRY(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } ev_guard{e0, e1};
    DIM_HIP_TRY(hipEventRecord(e0, stream));
    BState init{0, 0, static_cast<int>(o.max_iters), static_cast<int>(S), 0};
    DIM_HIP_TRY(hipMemcpyAsync(st, &init, sizeof(BState), hipMemcpyHostToDevice, stream));
    DIM_HIP_TRY(product(A, transpose, x, q, nullptr, stream));
    hipLaunchKernelGGL(k_b_init, dim3(B), dim3(kThreads), 0, stream, tb, b, static_cast<const double*>(q), r, p, rr,
                       bb);
    hipLaunchKernelGGL(k_b_init_state, dim3(1), dim3(64), 0, stream, st, sa, tb, static_cast<const double*>(rr),
                       static_cast<const double*>(bb), o.rtol);
    DIM_HIP_TRY(hipGetLastError());

    if (persistent) {
        // one cooperative launch runs every iteration; its vector work items
        // are the launch sequence's blocks (same rows, same reduction order)
        double *rr_part = nullptr, *dot_chunk = nullptr;
        unsigned* bar = nullptr;
        int* abort = nullptr;
        DIM_HIP_TRY(m.alloc(rr_part, B));
        DIM_HIP_TRY(m.alloc(dot_chunk, A->f.n_chunk_tab));
        // one block per CU when the kernel fits there (every block resident)
        const uint32_t grid = static_cast<uint32_t>(detail::device_cus());
        const size_t bar_words = size_t{kFlagStride} * (1 + grid);
        DIM_HIP_TRY(m.alloc(bar, bar_words));
        DIM_HIP_TRY(m.alloc(abort, 1));
        DIM_HIP_TRY(hipMemsetAsync(bar, 0, bar_words * sizeof(unsigned), stream));
        DIM_HIP_TRY(hipMemsetAsync(abort, 0, sizeof(int), stream));
        int dev = 0, rate_khz = 0;
        DIM_HIP_TRY(hipGetDevice(&dev));
        DIM_HIP_TRY(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
        PersistParams P{};
        P.tp = detail::make_product_params(A->f, A->s, p, q);
        P.tp.dot_chunk = dot_chunk;
        P.tb = tb;
        P.sa = sa;
        P.sys_chunk = A->d_sys_chunk;
        P.x = x;
        P.r = r;
        P.p = p;
        P.q = q;
        P.rr_part = rr_part;
        P.dot_chunk = dot_chunk;
        P.fixup = A->f.chunk_tab_fixup ? 1 : 0;
        P.nchunks = A->f.n_chunk_tab;
        P.n_items = B;
        P.n_sys = S;
        P.max_iters = static_cast<int>(o.max_iters);
        P.st = st;
        P.bar = bar;
        P.abort = abort;
        // a barrier waits at most 10 s (every wait is microseconds)
        P.bar_limit = uint64_t(std::max(rate_khz, 1)) * 1000ull * 10ull;
        // the initial state decides whether anything is left to iterate
        DIM_HIP_TRY(hipMemcpyAsync(hs, st, sizeof(BState), hipMemcpyDeviceToHost, stream));
        DIM_HIP_TRY(hipStreamSynchronize(stream));
        if (!hs->done) {
            DIM_HIP_TRY(launch_persistent(A->f.sigma, codes, grid, P, stream));
            int h_abort = 0;
            DIM_HIP_TRY(hipMemcpyAsync(&h_abort, abort, sizeof(int), hipMemcpyDeviceToHost, stream));
            DIM_HIP_TRY(hipMemcpyAsync(hs, st, sizeof(BState), hipMemcpyDeviceToHost, stream));
            DIM_HIP_TRY(hipStreamSynchronize(stream));
            if (h_abort)
                return fail(DIM_HIP_ERROR, "batched CG: a grid barrier of the persistent kernel timed out");
        }
    } else {
        // capture a batch of iterations, or reuse the handle's capture of the
        // same iterations
        auto enqueue = [&](hipStream_t s) {
            for (uint32_t it = 0; it < check; ++it) {
                const double* pq_parts = pq;
                const uint32_t* pq_range = d_sb;
                if (transpose) {
                    (void)product(A, true, p, q, &st->done, s);
                    hipLaunchKernelGGL(k_b_dot, dim3(B), dim3(kThreads), 0, s, st, sa, tb,
                                       static_cast<const double*>(p), static_cast<const double*>(q), pq);
                } else {
                    // p.q per chunk from the product's epilogue
                    (void)launch_spmv(A->f, A->s, false, 1, p, q, 0, A->f.tiles, 0, 0, A->f.n_empty, true, &st->done,
                                      s, nullptr, 0, 0, &pq_parts, sa.sys_done);
                    pq_range = A->d_sys_chunk;
                }
                hipLaunchKernelGGL(k_b_update_r, dim3(B), dim3(kThreads), 0, s, st, sa, tb, r,
                                   static_cast<const double*>(q), pq_parts, pq_range, rr);
                hipLaunchKernelGGL(k_b_update_xp, dim3(B), dim3(kThreads), 0, s, st, sa, tb, x,
                                   static_cast<const double*>(r), p, static_cast<const double*>(rr));
            }
        };
        hipGraphExec_t exec = nullptr;
        if (bc->exec && bc->g_transpose == transpose && bc->g_check == check && bc->g_x == x) {
            exec = bc->exec;
        } else {
            bc->reset_graph();
            bool ok = hipStreamBeginCapture(bc->cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (ok) {
                enqueue(bc->cap);
                ok = hipStreamEndCapture(bc->cap, &bc->graph) == hipSuccess && bc->graph &&
                     hipGraphInstantiate(&bc->exec, bc->graph, nullptr, nullptr, 0) == hipSuccess;
                (void)hipGetLastError();
            }
            if (ok) {
                exec = bc->exec;
                bc->g_transpose = transpose;
                bc->g_check = check;
                bc->g_x = x;
            } else {
                bc->reset_graph();
            }
        }
        const bool use_graph = exec != nullptr;
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
    }
    DIM_HIP_TRY(hipEventRecord(e1, stream));
    DIM_HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return report(A, b, x, transpose, tb, sys_blk, B, q, rt, sa, kThreads, *hs, ms, res, stream);
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
    if (opts->flags & ~(DIM_CG_TRANSPOSE | DIM_CG_PERSISTENT))
        return fail(DIM_INVALID_ARGUMENT, "unknown CG flags");
    if ((opts->flags & DIM_CG_TRANSPOSE) && (opts->flags & DIM_CG_PERSISTENT))
        return fail(DIM_INVALID_ARGUMENT, "the persistent batched CG does not iterate with A^T");
    return solve_batched(const_cast<dim_csr5*>(A), B, X, *opts, res, static_cast<hipStream_t>(stream));
}

} // extern "C"
