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
#pragma once

#include "dim_internal.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace dimgpu {
namespace detail {

template<int SIGMA>
struct DescWord {
    using type = uint32_t;
};
template<>
struct DescWord<8> {
    using type = uint16_t;
};
template<>
struct DescWord<16> {
    using type = uint16_t;
};

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD works on
// one contiguous eighth of the tiles and its L2 sees neighbouring x entries.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
    constexpr uint32_t X = 8;
    const uint32_t q = G / X, r = G % X, x = b % X, i = b / X;
    return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d)
            v += u;
    }
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

template<typename V>
__device__ __forceinline__ V ldg_stream(const V* p) {
    return __builtin_nontemporal_load(p);
}

struct TileParams {
    const uint32_t* tile_ptr;
    const void* bit_flag;
    const uint8_t* tile_flags;
    const uint32_t* tdo_ptr;
    const uint32_t* tdo;
    const uint32_t* col_idx;
    const void* vals;
    const void* x;
    void* y;
    void* head;
    void* tail;
    double* dot_tile;
    const int* done;
    uint64_t nnz;
    uint32_t tile_begin, tile_end, row_base, last_tile;
    int k;
};

template<typename V, int SIGMA, int KMAX, bool DOT>
__global__ __launch_bounds__(256) void k_csr5_tiles(TileParams p) {
    if (p.done && *p.done)
        return;
    const int lane = threadIdx.x & 63;
    const uint32_t t = p.tile_begin + xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (t >= p.tile_end)
        return; // wave-uniform

    using W = typename DescWord<SIGMA>::type;
    const int k = KMAX == 1 ? 1 : p.k;
    const V* __restrict__ x = static_cast<const V*>(p.x);
    V* __restrict__ y = static_cast<V*>(p.y);
    V* __restrict__ head = static_cast<V*>(p.head);
    V* __restrict__ tail = static_cast<V*>(p.tail);

    const size_t base = static_cast<size_t>(t) * (64 * SIGMA);
    const V* vals = static_cast<const V*>(p.vals) + base + lane;
    const uint32_t* cols = p.col_idx + base + lane;

    // Single-vector products keep the whole lane column in registers so that
    // all 2*sigma loads are in flight at once; multi-RHS products stream the
    // column step by step (their k-wide x rows dominate the register budget).
    constexpr int NREG = KMAX == 1 ? SIGMA : 1;
    V v[NREG];
    uint32_t c[NREG];
    if constexpr (KMAX == 1) {
#pragma unroll
        for (int r = 0; r < SIGMA; ++r) {
            v[r] = ldg_stream(vals + r * 64);
            c[r] = ldg_stream(cols + r * 64);
        }
    }

    const uint32_t bits = static_cast<uint32_t>(static_cast<const W*>(p.bit_flag)[static_cast<size_t>(t) * 64 + lane]) |
                          (lane == 0 ? 1u : 0u);
    const uint8_t tf = p.tile_flags[t];
    const bool starts_row = tf & kStartsRow;
    const uint32_t tpw = p.tile_ptr[t];
    const bool dirty = (tpw & kDirty) != 0;
    const uint32_t row0 = tpw & ~kDirty;
    const uint32_t* tdo = dirty ? p.tdo + p.tdo_ptr[t] : nullptr;

    // padded elements of the last tile contribute exactly zero
    uint32_t pad = 0;
    if (t == p.last_tile) {
#pragma unroll
        for (int r = 0; r < SIGMA; ++r)
            if (base + static_cast<size_t>(r) * 64 + lane >= p.nnz)
                pad |= 1u << r;
    }

    const int nb = __popc(bits);
    const int incl = wave_incl_scan(nb, lane);
    const int y_off = incl - nb;      // index of the first segment that starts in this lane
    const int nseg = __shfl(incl, 63, 64);
    const uint64_t started = __ballot(nb > 0);

    V cur[KMAX], red[KMAX];
    double dacc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        cur[j] = V(0);
        red[j] = V(0);
        dacc[j] = 0.0;
    }

    auto emit = [&](int s, const V* val) {
        if (s == 0 && !starts_row) {
            // the tile's first segment continues a row from an earlier tile
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                    head[static_cast<size_t>(t) * k + j] = val[j];
            return;
        }
        const uint32_t row = row0 + (dirty ? tdo[s] : static_cast<uint32_t>(s));
        V* yr = y + static_cast<size_t>(row - p.row_base) * k;
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            if (j < k)
                yr[j] = val[j];
        if constexpr (DOT) {
            const V* pr = x + static_cast<size_t>(row) * k;
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                    dacc[j] += static_cast<double>(pr[j]) * static_cast<double>(val[j]);
        }
    };

    bool has = false;
    int seg = y_off - 1;
    auto step = [&](int r, V vr, uint32_t cr) {
        if ((bits >> r) & 1u) {
            if (has)
                emit(seg, cur);
            else
#pragma unroll
                for (int j = 0; j < KMAX; ++j)
                    red[j] = cur[j];
            has = true;
            ++seg;
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                cur[j] = V(0);
        }
        const V* xr = x + static_cast<size_t>(cr) * k;
        const bool pz = (pad >> r) & 1u;
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            if (j < k) {
                V xv = xr[j];
                if (pz)
                    xv = V(0);
                cur[j] = fma(vr, xv, cur[j]);
            }
        }
    };
    if constexpr (KMAX == 1) {
#pragma unroll
        for (int r = 0; r < SIGMA; ++r)
            step(r, v[r], c[r]);
    } else {
#pragma unroll 1
        for (int r = 0; r < SIGMA; ++r)
            step(r, ldg_stream(vals + r * 64), ldg_stream(cols + r * 64));
    }

    V blue[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (has) {
            blue[j] = cur[j];
        } else {
            red[j] = cur[j];
            blue[j] = V(0);
        }
    }

    // Segmented inclusive scan of the lanes' leading (red) partials; a new
    // segment begins after every lane in which a row starts.
    int f = (lane == 0) ? 1 : static_cast<int>((started >> (lane - 1)) & 1ull);
    V A[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        A[j] = red[j];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int fup = __shfl_up(f, d, 64);
        V up[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            up[j] = __shfl_up(A[j], d, 64);
        if (lane >= d && !f) {
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                A[j] = up[j] + A[j];
            f = fup;
        }
    }

    // A lane in which a row starts closes the row left open by the previous
    // such lane: its trailing (blue) partial plus the scanned red partials.
    const uint64_t below = lane ? (started & ((1ull << lane) - 1ull)) : 0ull;
    const int cprev = below ? 63 - __clzll(below) : 0;
    V bc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        bc[j] = __shfl(blue[j], cprev, 64);
    if (lane > 0 && has) {
        V tot[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            tot[j] = bc[j] + A[j];
        emit(y_off - 1, tot);
    }

    // The tile's last segment.
    const int L = 63 - __clzll(started);
    V bl[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        bl[j] = __shfl(blue[j], L, 64);
    if (lane == 63) {
        V tv[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            tv[j] = (L == 63) ? blue[j] : bl[j] + A[j];
        const bool complete = (t + 1 == p.tile_end) || (p.tile_flags[t + 1] & kStartsRow);
        if (nseg == 1 && !starts_row) {
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k) {
                    head[static_cast<size_t>(t) * k + j] = tv[j];
                    tail[static_cast<size_t>(t) * k + j] = tv[j];
                }
        } else if (complete) {
            emit(nseg - 1, tv);
        } else {
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                    tail[static_cast<size_t>(t) * k + j] = tv[j];
        }
    }

    if constexpr (DOT) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const double s = wave_sum(dacc[j]);
            if (lane == 0 && j < k)
                p.dot_tile[static_cast<size_t>(t) * k + j] = s;
        }
    }
}

struct FixParams {
    const uint32_t* tile_ptr;
    const uint8_t* tile_flags;
    const uint32_t* empty_rows;
    const void* x;
    void* y;
    const void* head;
    const void* tail;
    const double* dot_tile;
    double* dot_blk;
    const int* done;
    uint32_t tile_begin, tile_end, row_base, empty_begin, empty_end;
    int k;
};

// Rows that cross tile boundaries: the tile in which such a row ends adds the
// chain tail(start) + tail(middle tiles...) + head(end) in tile order and
// writes the row once. Also writes 0 to empty rows and, for CG, reduces the
// per-tile dot partials of 256 tiles to one partial per block.
template<typename V, int KMAX, bool DOT>
__global__ __launch_bounds__(256) void k_csr5_fixup(FixParams p) {
    if (p.done && *p.done)
        return;
    const uint32_t t = p.tile_begin + blockIdx.x * 256 + threadIdx.x;
    const int k = KMAX == 1 ? 1 : p.k;
    const V* x = static_cast<const V*>(p.x);
    V* y = static_cast<V*>(p.y);
    const V* head = static_cast<const V*>(p.head);
    const V* tail = static_cast<const V*>(p.tail);
    double dacc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        dacc[j] = 0.0;

    if (t < p.tile_end) {
        const uint8_t tf = p.tile_flags[t];
        if (!(tf & kStartsRow)) {
            const bool closing =
              (tf & kMultiSeg) || (t + 1 == p.tile_end) || (p.tile_flags[t + 1] & kStartsRow);
            if (closing) {
                int64_t kk = static_cast<int64_t>(t) - 1;
                while (kk >= static_cast<int64_t>(p.tile_begin) && !(p.tile_flags[kk] & (kStartsRow | kMultiSeg)))
                    --kk;
                const uint32_t from = kk >= static_cast<int64_t>(p.tile_begin) ? static_cast<uint32_t>(kk) : p.tile_begin;
                const uint32_t row = p.tile_ptr[t] & ~kDirty;
                V* yr = y + static_cast<size_t>(row - p.row_base) * k;
#pragma unroll
                for (int j = 0; j < KMAX; ++j) {
                    if (j >= k)
                        break;
                    V s;
                    if (from < t) {
                        s = tail[static_cast<size_t>(from) * k + j];
                        for (uint32_t u = from + 1; u < t; ++u)
                            s = s + tail[static_cast<size_t>(u) * k + j];
                        s = s + head[static_cast<size_t>(t) * k + j];
                    } else {
                        s = head[static_cast<size_t>(t) * k + j];
                    }
                    yr[j] = s;
                    if constexpr (DOT)
                        dacc[j] += static_cast<double>(x[static_cast<size_t>(row) * k + j]) * static_cast<double>(s);
                }
            }
        }
        if constexpr (DOT) {
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                    dacc[j] += p.dot_tile[static_cast<size_t>(t) * k + j];
        }
    }

    const uint32_t ne = p.empty_end - p.empty_begin;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ne; i += gridDim.x * 256) {
        const uint32_t row = p.empty_rows[p.empty_begin + i];
        V* yr = y + static_cast<size_t>(row - p.row_base) * k;
        for (int j = 0; j < k; ++j)
            yr[j] = V(0);
    }

    if constexpr (DOT) {
        __shared__ double red[4][KMAX];
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const double s = wave_sum(dacc[j]);
            if (lane == 0)
                red[w][j] = s;
        }
        __syncthreads();
        if (threadIdx.x < static_cast<unsigned>(k))
            p.dot_blk[static_cast<size_t>(blockIdx.x) * k + threadIdx.x] =
              (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    }
}

template<typename V, int SIGMA, int KMAX, bool DOT>
void launch_tiles(const TileParams& tp, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL((k_csr5_tiles<V, SIGMA, KMAX, DOT>), dim3(grid), dim3(256), 0, st, tp);
}

template<typename V, int KMAX, bool DOT>
void launch_by_sigma(uint32_t sigma, const TileParams& tp, uint32_t grid, hipStream_t st) {
    switch (sigma) {
        case 8: launch_tiles<V, 8, KMAX, DOT>(tp, grid, st); break;
        case 16: launch_tiles<V, 16, KMAX, DOT>(tp, grid, st); break;
        case 24: launch_tiles<V, 24, KMAX, DOT>(tp, grid, st); break;
        default: launch_tiles<V, 32, KMAX, DOT>(tp, grid, st); break;
    }
}

template<typename V, int KMAX>
void launch_pair(const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid, uint32_t fgrid,
                 bool dot, hipStream_t st) {
    if (tgrid) {
        if (dot)
            launch_by_sigma<V, KMAX, true>(A.sigma, tp, tgrid, st);
        else
            launch_by_sigma<V, KMAX, false>(A.sigma, tp, tgrid, st);
    }
    if (dot)
        hipLaunchKernelGGL((k_csr5_fixup<V, KMAX, true>), dim3(fgrid), dim3(256), 0, st, fp);
    else
        hipLaunchKernelGGL((k_csr5_fixup<V, KMAX, false>), dim3(fgrid), dim3(256), 0, st, fp);
}

// Multi-RHS launchers, one translation unit per k bound (spmv_k*.hip).
void launch_pair_k2(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid,
                    uint32_t fgrid, bool dot, hipStream_t st);
void launch_pair_k4(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid,
                    uint32_t fgrid, bool dot, hipStream_t st);
void launch_pair_k8(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid,
                    uint32_t fgrid, bool dot, hipStream_t st);
void launch_pair_k16(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid,
                     uint32_t fgrid, bool dot, hipStream_t st);

} // namespace detail
} // namespace dimgpu
