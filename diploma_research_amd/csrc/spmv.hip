// CSR5 SpMV / SpMM kernels for gfx950 (CDNA4, wave64).
//
// One wavefront owns one tile of 64 columns x sigma steps. The reference does
// the same work with four AVX2 lanes per tile and OpenMP threads over tile
// chunks (include/dim/mat/storage_formats/csr5.h:619-748); here the 64 tile
// columns are the 64 lanes, so:
//   * step r of the tile is one coalesced load of 64 values and 64 column
//     indices (the per-tile transpose of csr5.h:463-493 puts them side by side);
//   * the lane's bit flags are decoded in registers, the per-column y_offset is
//     a wavefront prefix sum of popcounts and scansum_offset is read off a
//     ballot of the lanes that start a row (csr5.h:74-97);
//   * rows that cross lanes are closed with a segmented wavefront scan of the
//     lanes' leading partial sums, which replaces fast_segmented_sum /
//     compute_last_sum (csr5.h:200-235) and keeps each sum in left-to-right
//     order instead of differencing prefix sums;
//   * rows that cross tiles leave their partial sums in per-tile head/tail
//     slots, and a second kernel adds the chains up in tile order and writes
//     those rows once. This replaces the reference's per-thread calibrator
//     (csr5.h:526-538, 750-754) without atomics, so results are bitwise
//     reproducible from run to run.
#include "spmv_impl.h"

#include <algorithm>

namespace dimgpu {

namespace {

using namespace detail;

// ------------------------------------------------------ transpose SpMV ---

template<typename V>
__global__ void k_fill(V* y, size_t n, V value, const int* done) {
    if (done && *done)
        return;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        y[i] = value;
}

// Row of the element at step r of this lane, from the lane's bit flags.
__device__ __forceinline__ uint32_t element_row(uint32_t bits, int r, int y_off, uint32_t row0, bool dirty,
                                                const uint32_t* tdo) {
    const uint32_t upto = (r >= 31) ? 0xFFFFFFFFu : ((2u << r) - 1u);
    const int s = y_off - 1 + __popc(bits & upto);
    return row0 + (dirty ? tdo[s] : static_cast<uint32_t>(s));
}

// y[col] += a_ij * x[row] for every element: the transposed product scatters
// into y with hardware floating-point atomics.
template<typename V, int SIGMA>
__global__ __launch_bounds__(256) void k_csr5_transpose(TileParams p) {
    if (p.done && *p.done)
        return;
    const int lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= p.tile_end)
        return;
    using W = typename DescWord<SIGMA>::type;
    const V* x = static_cast<const V*>(p.x);
    V* y = static_cast<V*>(p.y);
    const size_t base = static_cast<size_t>(t) * (64 * SIGMA);
    const uint32_t bits = static_cast<uint32_t>(static_cast<const W*>(p.bit_flag)[static_cast<size_t>(t) * 64 + lane]) |
                          (lane == 0 ? 1u : 0u);
    const uint32_t tpw = p.tile_ptr[t];
    const bool dirty = (tpw & kDirty) != 0;
    const uint32_t row0 = tpw & ~kDirty;
    const uint32_t* tdo = dirty ? p.tdo + p.tdo_ptr[t] : nullptr;
    const int nb = __popc(bits);
    const int y_off = wave_incl_scan(nb, lane) - nb;
    const V* vals = static_cast<const V*>(p.vals);
#pragma unroll
    for (int r = 0; r < SIGMA; ++r) {
        const size_t pos = base + static_cast<size_t>(r) * 64 + lane;
        const size_t csr = base + static_cast<size_t>(lane) * SIGMA + r;
        if (csr >= p.nnz)
            continue;
        const uint32_t row = element_row(bits, r, y_off, row0, dirty, tdo);
        const V a = vals[pos];
        unsafeAtomicAdd(y + p.col_idx[pos], a * x[row]);
    }
}

// out[csr(e)] = scale * u[row(e)] * v[col(e)]
template<int SIGMA>
__global__ __launch_bounds__(256) void k_sampled_outer(TileParams p, double scale, const double* u, const double* v,
                                                       double* out) {
    const int lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= p.tile_end)
        return;
    using W = typename DescWord<SIGMA>::type;
    const size_t base = static_cast<size_t>(t) * (64 * SIGMA);
    const uint32_t bits = static_cast<uint32_t>(static_cast<const W*>(p.bit_flag)[static_cast<size_t>(t) * 64 + lane]) |
                          (lane == 0 ? 1u : 0u);
    const uint32_t tpw = p.tile_ptr[t];
    const bool dirty = (tpw & kDirty) != 0;
    const uint32_t row0 = tpw & ~kDirty;
    const uint32_t* tdo = dirty ? p.tdo + p.tdo_ptr[t] : nullptr;
    const int nb = __popc(bits);
    const int y_off = wave_incl_scan(nb, lane) - nb;
#pragma unroll
    for (int r = 0; r < SIGMA; ++r) {
        const size_t pos = base + static_cast<size_t>(r) * 64 + lane;
        const size_t csr = base + static_cast<size_t>(lane) * SIGMA + r;
        if (csr >= p.nnz)
            continue;
        const uint32_t row = element_row(bits, r, y_off, row0, dirty, tdo);
        out[csr] = scale * u[row] * v[p.col_idx[pos]];
    }
}

// Tile-layout values from CSR-order values: the same permutation as the
// conversion's transpose, run on the device.
template<int SIGMA>
__global__ __launch_bounds__(256) void k_set_values(TileParams p, const double* src, double* v64, float* v32) {
    const int lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= p.tile_end)
        return;
    const size_t base = static_cast<size_t>(t) * (64 * SIGMA);
#pragma unroll
    for (int r = 0; r < SIGMA; ++r) {
        const size_t pos = base + static_cast<size_t>(r) * 64 + lane;
        const size_t csr = base + static_cast<size_t>(lane) * SIGMA + r;
        const double a = csr < p.nnz ? src[csr] : 0.0;
        v64[pos] = a;
        if (v32)
            v32[pos] = static_cast<float>(a);
    }
}

TileParams make_tile_params(const DevFormat& A, bool f32) {
    TileParams tp{};
    tp.tile_ptr = A.tile_ptr;
    tp.bit_flag = A.bit_flag;
    tp.tile_flags = A.tile_flags;
    tp.tdo_ptr = A.tile_desc_offset_ptr;
    tp.tdo = A.tile_desc_offset;
    tp.col_idx = A.col_idx;
    tp.col_code = A.col_code;
    tp.col_dict = A.col_dict;
    tp.cols = A.cols;
    tp.vals = f32 ? static_cast<const void*>(A.vals32) : static_cast<const void*>(A.vals64);
    tp.nnz = A.nnz;
    tp.tile_begin = 0;
    tp.tile_end = A.tiles;
    tp.last_tile = A.tiles ? A.tiles - 1 : 0;
    tp.k = 1;
    return tp;
}

} // namespace

using namespace detail;

bool BoundaryInfo::needs_fixup(uint32_t tb, uint32_t te, uint32_t tpw) const {
    std::lock_guard<std::mutex> g(m_);
    if (memo_valid_ && memo_tb_ == tb && memo_te_ == te && memo_tpw_ == tpw)
        return memo_need_;
    bool need = te > flags_.size() || tpw == 0;
    if (!need && tb < te) {
        // the first chunk keeps a row that began before the range in its head slot
        need = !(flags_[tb] & kStartsRow);
        for (uint64_t t = uint64_t{tb} + tpw; !need && t < te; t += tpw) {
            const uint8_t f = flags_[t];
            const bool looked = (f & kMultiSeg) && (flags_[t - 1] & (kStartsRow | kMultiSeg));
            need = !(f & kStartsRow) && !looked;
        }
    }
    memo_valid_ = true;
    memo_tb_ = tb;
    memo_te_ = te;
    memo_tpw_ = tpw;
    memo_need_ = need;
    return need;
}

bool BoundaryInfo::needs_fixup_table(const std::vector<uint32_t>& chunk_tile) const {
    std::lock_guard<std::mutex> g(m_);
    if (chunk_tile.empty() || chunk_tile.back() > flags_.size())
        return true;
    bool need = chunk_tile.size() > 1 && !(flags_[chunk_tile[0]] & kStartsRow);
    for (size_t c = 1; !need && c + 1 < chunk_tile.size(); ++c) {
        const uint32_t t = chunk_tile[c];
        const uint8_t f = flags_[t];
        const bool looked = (f & kMultiSeg) && (flags_[t - 1] & (kStartsRow | kMultiSeg));
        need = !(f & kStartsRow) && !looked;
    }
    return need;
}

// Measured on the MI355X, FP64 SpMM of the 192^3 variable-coefficient
// operator (sigma 7, column codes), k = 8: one chunk per wavefront 675 us;
// ordered sweep with chunks of 1, 2, 4, 8, 16 tiles: 542, 514, 502, 534,
// 565 us (k = 2: 165 us either way).
uint32_t detail::sweep_tiles() {
    return 4;
}

bool detail::solve_streams_matrix(const DevFormat& A, size_t vb) {
    constexpr uint64_t kInfinityCacheBytes = 256ull << 20;
    const uint64_t lanes = static_cast<uint64_t>(A.tiles) * 64;
    const uint64_t elems = lanes * A.sigma;
    uint64_t bytes = elems * vb; // values
    if (A.col_code)
        bytes += lanes * 4; // one code word per lane, flags included
    else
        bytes += elems * 4 + lanes * (A.sigma <= 16 ? 2 : 4); // column indices, row-start flags
    bytes += 4ull * A.rows * vb;                               // x, r, p, q
    return bytes > kInfinityCacheBytes;
}

int detail::device_cus() {
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return 256;
    if (!cus[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

namespace {

// widest multi-RHS block decoded in one pass over the tiles
constexpr int kColumnGroupF64 = 16;
constexpr int kColumnGroupF32 = 16;

void fill_params(const DevFormat& A, Scratch& S, bool f32, int k, const void* x, void* y, uint32_t tile_begin,
                 uint32_t tile_end, uint32_t row_base, uint32_t empty_begin, uint32_t empty_end, const int* done,
                 TileParams& tp, FixParams& fp, uint64_t ldx = 0, uint64_t ldy = 0) {
    tp = make_tile_params(A, f32);
    tp.x = x;
    tp.y = y;
    tp.head = S.head;
    tp.tail = S.tail;
    tp.chunk_flags = S.chunk_flags;
    tp.dot_chunk = S.dot_chunk;
    tp.queue = S.queue;
    tp.done = done;
    tp.tile_begin = tile_begin;
    tp.tile_end = tile_end;
    tp.row_base = row_base;
    tp.k = k;
    tp.empty_rows = A.empty_rows;
    tp.empty_begin = empty_begin;
    tp.empty_end = empty_end;
    tp.rs = ldx ? 1 : static_cast<uint64_t>(k);
    tp.xcs = ldx ? ldx : 1;
    tp.ycs = ldx ? (ldy ? ldy : ldx) : 1;
    fp = FixParams{};
    fp.rs = tp.rs;
    fp.xcs = tp.xcs;
    fp.ycs = tp.ycs;
    fp.dot_ld = static_cast<uint32_t>(k);
    fp.dot_off = 0;
    fp.tile_ptr = A.tile_ptr;
    fp.empty_rows = A.empty_rows;
    fp.chunk_flags = S.chunk_flags;
    fp.x = x;
    fp.y = y;
    fp.head = S.head;
    fp.tail = S.tail;
    fp.dot_chunk = S.dot_chunk;
    fp.dot_blk = S.dot_blk;
    fp.done = done;
    fp.tile_begin = tile_begin;
    fp.row_base = row_base;
    fp.empty_begin = empty_begin;
    fp.empty_end = empty_end;
    fp.k = k;
}

} // namespace

uint64_t detail::resident_tile_waves(uint32_t sigma, bool codes, bool dot) {
    int b = 2;
    switch (sigma) {
#define DIM_WAVES_CASE(S)                                                                                             \
    case S:                                                                                                           \
        if constexpr (coded_kernel<1, S>())                                                                           \
            if (codes) {                                                                                              \
                b = dot ? tile_blocks_per_cu<double, S, 1, true, true>() : tile_blocks_per_cu<double, S, 1, false, true>(); \
                break;                                                                                                \
            }                                                                                                         \
        b = dot ? tile_blocks_per_cu<double, S, 1, true, false>() : tile_blocks_per_cu<double, S, 1, false, false>();  \
        break;
        DIM_FOR_EACH_SIGMA(DIM_WAVES_CASE)
#undef DIM_WAVES_CASE
        default: break;
    }
    return uint64_t(device_cus()) * b * 4;
}

TileParams detail::make_product_params(const DevFormat& A, Scratch& S, const void* x, void* y) {
    TileParams tp;
    FixParams fp;
    fill_params(A, S, false, 1, x, y, 0, A.tiles, 0, 0, A.n_empty, nullptr, tp, fp);
    tp.chunk_tile = A.chunk_tile;
    tp.chunk_sys = A.chunk_sys;
    tp.nchunks = A.n_chunk_tab;
    tp.tpw = 0;
    tp.lookahead = 1;
    return tp;
}

hipError_t launch_spmv(const DevFormat& A, Scratch& S, bool f32, int k, const void* x, void* y, uint32_t tile_begin,
                       uint32_t tile_end, uint32_t row_base, uint32_t empty_begin, uint32_t empty_end, bool dot,
                       const int* done, hipStream_t stream, uint32_t* dot_blocks, uint64_t ldx, uint64_t ldy,
                       const double** dot_parts, const int* sys_done) {
    TileParams tp;
    FixParams fp;
    fill_params(A, S, f32, k, x, y, tile_begin, tile_end, row_base, empty_begin, empty_end, done, tp, fp, ldx, ldy);
    if (k == 1)
        tp.sys_done = sys_done;
    uint32_t nb = 0;
    const double* parts = S.dot_blk;
    if (k == 1) {
        nb = f32 ? launch_pair<float, 1>(A, tp, fp, dot, stream, nullptr, &parts)
                 : launch_pair<double, 1>(A, tp, fp, dot, stream, nullptr, &parts);
    } else {
        // Wide blocks may run as column groups over the same tiles (each group
        // reads the matrix again). Every group runs the same kernel width, so
        // that all groups share one launch plan and leave the same number of
        // dot partials per column.
        const int gmax = f32 ? kColumnGroupF32 : kColumnGroupF64;
        const int group = k < gmax ? k : gmax;
        const size_t esz = f32 ? 4 : 8;
        for (int j0 = 0; j0 < k; j0 += group) {
            const int kc = k - j0 < group ? k - j0 : group;
            TileParams t = tp;
            FixParams f = fp;
            t.k = f.k = kc;
            t.x = f.x = static_cast<const char*>(x) + j0 * tp.xcs * esz;
            t.y = f.y = static_cast<char*>(y) + j0 * tp.ycs * esz;
            f.dot_off = static_cast<uint32_t>(j0);
            if (group <= 2)
                nb = launch_pair_k2(f32, A, t, f, dot, stream);
            else if (group <= 4)
                nb = launch_pair_k4(f32, A, t, f, dot, stream);
            else if (group <= 8)
                nb = launch_pair_k8(f32, A, t, f, dot, stream);
            else
                nb = launch_pair_k16(f32, A, t, f, dot, stream);
        }
    }
    if (dot_blocks)
        *dot_blocks = nb;
    if (dot_parts)
        *dot_parts = parts;
    return hipGetLastError();
}

hipError_t launch_spmv_events(const DevFormat& A, Scratch& S, bool f32, const void* x, void* y, hipEvent_t e0,
                              hipEvent_t e1, hipEvent_t e2, hipStream_t stream, bool* fixup_launched) {
    TileParams tp;
    FixParams fp;
    fill_params(A, S, f32, 1, x, y, 0, A.tiles, 0, 0, A.n_empty, nullptr, tp, fp);
    hipError_t e = hipEventRecord(e0, stream);
    if (e != hipSuccess)
        return e;
    const uint32_t nfix =
      f32 ? launch_pair<float, 1>(A, tp, fp, false, stream, e1) : launch_pair<double, 1>(A, tp, fp, false, stream, e1);
    if (fixup_launched)
        *fixup_launched = nfix != 0;
    e = hipEventRecord(e2, stream);
    if (e != hipSuccess)
        return e;
    return hipGetLastError();
}

namespace {

// One row per thread; the S slices' bit words, ranks and partials are each
// loaded together (all in flight at once) before the in-order sum.
template<typename V, int S>
__global__ __launch_bounds__(256) void k_slice_sum(const V* __restrict__ ys, const uint32_t* __restrict__ present,
                                                   const uint32_t* __restrict__ rank, uint32_t rows,
                                                   V* __restrict__ y) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= rows)
        return;
    uint32_t w[S], at[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const size_t r = static_cast<size_t>(q) * rows + i;
        w[q] = present[r >> 5];
        at[q] = rank[r >> 5];
    }
    V p[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
        const uint32_t b = static_cast<uint32_t>((static_cast<size_t>(q) * rows + i) & 31u);
        const bool has = (w[q] >> b) & 1u;
        p[q] = has ? ys[at[q] + __popc(w[q] & ((1u << b) - 1u))] : V(0);
    }
    V s = V(0);
#pragma unroll
    for (int q = 0; q < S; ++q)
        if ((w[q] >> ((static_cast<size_t>(q) * rows + i) & 31u)) & 1u)
            s = s + p[q];
    y[i] = s;
}

} // namespace

namespace {

// n x k blocks between row-major (i*k + j) and column-major (j*n + i),
// k <= 16: a block moves 256 rows through LDS, so that both the row-major
// and the column-major side are read and written in contiguous runs.
constexpr int kTransposeRows = 256;
template<typename V, bool TO_COLS>
__global__ __launch_bounds__(256) void k_block_transpose(size_t n, int k, const V* __restrict__ src,
                                                         V* __restrict__ dst) {
    __shared__ V tile[16 * (kTransposeRows + 1)];
    const size_t row0 = static_cast<size_t>(blockIdx.x) * kTransposeRows;
    const int rows = static_cast<int>(n - row0 < kTransposeRows ? n - row0 : kTransposeRows);
    const int count = rows * k;
    if constexpr (TO_COLS) {
        const V* s = src + row0 * k;
        for (int e = threadIdx.x; e < count; e += 256)
            tile[(e % k) * (kTransposeRows + 1) + e / k] = s[e];
        __syncthreads();
        for (int e = threadIdx.x; e < count; e += 256) {
            const int j = e / rows, r = e - j * rows;
            dst[static_cast<size_t>(j) * n + row0 + r] = tile[j * (kTransposeRows + 1) + r];
        }
    } else {
        for (int e = threadIdx.x; e < count; e += 256) {
            const int j = e / rows, r = e - j * rows;
            tile[j * (kTransposeRows + 1) + r] = src[static_cast<size_t>(j) * n + row0 + r];
        }
        __syncthreads();
        V* d = dst + row0 * k;
        for (int e = threadIdx.x; e < count; e += 256)
            d[e] = tile[(e % k) * (kTransposeRows + 1) + e / k];
    }
}

} // namespace

hipError_t launch_block_transpose(bool f32, bool to_cols, size_t n, int k, const void* src, void* dst,
                                  hipStream_t stream) {
    if (!n)
        return hipSuccess;
    if (k < 1 || k > 16)
        return hipErrorInvalidValue;
    const uint32_t grid = static_cast<uint32_t>((n + kTransposeRows - 1) / kTransposeRows);
    if (f32) {
        if (to_cols)
            hipLaunchKernelGGL((k_block_transpose<float, true>), dim3(grid), dim3(256), 0, stream, n, k,
                               static_cast<const float*>(src), static_cast<float*>(dst));
        else
            hipLaunchKernelGGL((k_block_transpose<float, false>), dim3(grid), dim3(256), 0, stream, n, k,
                               static_cast<const float*>(src), static_cast<float*>(dst));
    } else {
        if (to_cols)
            hipLaunchKernelGGL((k_block_transpose<double, true>), dim3(grid), dim3(256), 0, stream, n, k,
                               static_cast<const double*>(src), static_cast<double*>(dst));
        else
            hipLaunchKernelGGL((k_block_transpose<double, false>), dim3(grid), dim3(256), 0, stream, n, k,
                               static_cast<const double*>(src), static_cast<double*>(dst));
    }
    return hipGetLastError();
}

hipError_t launch_slice_sum(bool f32, const void* ystack, const uint32_t* present, const uint32_t* rank,
                            uint32_t rows, uint32_t slices, void* y, hipStream_t stream) {
    if (!rows)
        return hipSuccess;
    if (slices != 8)
        return hipErrorInvalidValue; // the library builds 8 slices (one per XCD)
    const uint32_t grid = (rows + 255) / 256;
    if (f32)
        hipLaunchKernelGGL((k_slice_sum<float, 8>), dim3(grid), dim3(256), 0, stream,
                           static_cast<const float*>(ystack), present, rank, rows, static_cast<float*>(y));
    else
        hipLaunchKernelGGL((k_slice_sum<double, 8>), dim3(grid), dim3(256), 0, stream,
                           static_cast<const double*>(ystack), present, rank, rows, static_cast<double*>(y));
    return hipGetLastError();
}

hipError_t launch_spmv_transpose(const DevFormat& A, bool f32, const void* x, void* y, hipStream_t stream,
                                 const int* done) {
    TileParams tp = make_tile_params(A, f32);
    tp.x = x;
    tp.y = y;
    tp.done = done;
    const uint32_t fill_grid = 1024;
    if (f32)
        hipLaunchKernelGGL(k_fill<float>, dim3(fill_grid), dim3(256), 0, stream, static_cast<float*>(y),
                           static_cast<size_t>(A.cols), 0.0f, done);
    else
        hipLaunchKernelGGL(k_fill<double>, dim3(fill_grid), dim3(256), 0, stream, static_cast<double*>(y),
                           static_cast<size_t>(A.cols), 0.0, done);
    if (A.tiles) {
        const uint32_t grid = (A.tiles + 3) / 4;
#define DIM_T_F(S)                                                                                                    \
    case S: hipLaunchKernelGGL((k_csr5_transpose<float, S>), dim3(grid), dim3(256), 0, stream, tp); break;
#define DIM_T_D(S)                                                                                                    \
    case S: hipLaunchKernelGGL((k_csr5_transpose<double, S>), dim3(grid), dim3(256), 0, stream, tp); break;
        if (f32) {
            switch (A.sigma) { DIM_FOR_EACH_SIGMA(DIM_T_F) default: break; }
        } else {
            switch (A.sigma) { DIM_FOR_EACH_SIGMA(DIM_T_D) default: break; }
        }
#undef DIM_T_F
#undef DIM_T_D
    }
    return hipGetLastError();
}

hipError_t launch_set_values(const DevFormat& A, const double* vals_csr, hipStream_t stream) {
    if (!A.tiles)
        return hipSuccess;
    TileParams tp = make_tile_params(A, false);
    const uint32_t grid = (A.tiles + 3) / 4;
    switch (A.sigma) {
#define DIM_SV(S)                                                                                                     \
    case S:                                                                                                           \
        hipLaunchKernelGGL((k_set_values<S>), dim3(grid), dim3(256), 0, stream, tp, vals_csr, A.vals64, A.vals32);   \
        break;
        DIM_FOR_EACH_SIGMA(DIM_SV)
#undef DIM_SV
        default: break;
    }
    return hipGetLastError();
}

hipError_t launch_sampled_outer(const DevFormat& A, double scale, const double* u, const double* v, double* out,
                                hipStream_t stream) {
    if (!A.tiles)
        return hipSuccess;
    TileParams tp = make_tile_params(A, false);
    const uint32_t grid = (A.tiles + 3) / 4;
    switch (A.sigma) {
#define DIM_O(S)                                                                                                      \
    case S: hipLaunchKernelGGL((k_sampled_outer<S>), dim3(grid), dim3(256), 0, stream, tp, scale, u, v, out); break;
        DIM_FOR_EACH_SIGMA(DIM_O)
#undef DIM_O
        default: break;
    }
    return hipGetLastError();
}

} // namespace dimgpu
