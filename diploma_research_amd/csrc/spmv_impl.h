// CSR5 SpMV / SpMM kernels for gfx950 (CDNA4, wave64).
//
// The reference runs CSR5 with four AVX2 lanes per tile and OpenMP threads
// over contiguous tile chunks (include/dim/mat/storage_formats/csr5.h:
// 608-748). Here a tile is 64 columns wide, one column per lane of a
// wavefront, and every wavefront walks one contiguous chunk of tiles:
//   * step r of a tile is one coalesced load of 64 values and 64 column
//     indices (the per-tile transpose of csr5.h:463-493 puts them side by side);
//   * the lane's bit flags are decoded in registers: the per-column y_offset is
//     a wavefront prefix sum of popcounts and scansum_offset is read off a
//     ballot of the lanes that start a row (csr5.h:74-97);
//   * rows that cross lanes are closed with a segmented wavefront scan of the
//     lanes' leading partial sums, which replaces fast_segmented_sum /
//     compute_last_sum (csr5.h:200-235) and adds each row left to right
//     instead of differencing prefix sums;
//   * a row that crosses from one tile into the next tile of the same chunk is
//     carried in registers; only rows that cross a chunk boundary leave
//     partial sums in per-chunk head/tail slots, and a small second kernel
//     adds those chains up in chunk order and writes each such row once. This
//     replaces the reference's per-thread calibrator (csr5.h:526-538,
//     750-754) without atomics, so results are bitwise reproducible.
// Chunks are sized so that all wavefronts of the launch are resident at once
// (one chunk per wavefront): a launch of one short wavefront per tile is
// limited by the rate at which workgroups are dispatched.
#pragma once

#include "dim_internal.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace dimgpu {
namespace detail {

// bit flags: one bit per step of a lane's column
template<int SIGMA>
struct DescWord {
    using type = typename std::conditional<(SIGMA <= 16), uint16_t, uint32_t>::type;
};

// Blocks are dealt round-robin to the 8 XCDs; remap so that each XCD works on
// one contiguous eighth of the chunks and its L2 sees neighbouring x entries.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
    constexpr uint32_t X = 8;
    const uint32_t q = G / X, r = G % X, x = b % X, i = b / X;
    return x * q + (x < r ? x : r) + i;
}

// ---- wavefront scans on DPP (data-parallel primitives: lane moves done by
// the VALU itself, no LDS round trip). Inclusive scans over the 64 lanes use
// row_shr 1,2,4,8 inside each row of 16 lanes, then row_bcast:15 and
// row_bcast:31 to carry row totals into the rows above.
constexpr int kDppRowShr1 = 0x111;
constexpr int kDppRowShr2 = 0x112;
constexpr int kDppRowShr4 = 0x114;
constexpr int kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;
constexpr int kDppWaveShr1 = 0x138; // lane i takes lane i-1 (lane 0 takes 0)

template<int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, true);
}
template<int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_val(float v) {
    return __int_as_float(dpp_int<CTRL, ROW_MASK>(__float_as_int(v)));
}
template<int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_val(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_int<CTRL, ROW_MASK>(static_cast<int>(b & 0xffffffffll));
    const int hi = dpp_int<CTRL, ROW_MASK>(static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template<int CTRL, int ROW_MASK>
__device__ __forceinline__ void scan_step(int& v) {
    v += dpp_int<CTRL, ROW_MASK>(v);
}

__device__ __forceinline__ int wave_incl_scan(int v, int) {
    scan_step<kDppRowShr1, 0xf>(v);
    scan_step<kDppRowShr2, 0xf>(v);
    scan_step<kDppRowShr4, 0xf>(v);
    scan_step<kDppRowShr8, 0xf>(v);
    scan_step<kDppRowBcast15, 0xa>(v);
    scan_step<kDppRowBcast31, 0xc>(v);
    return v;
}

// One step of a segmented inclusive scan: (a, f) = (up_a, up_f) (+) (a, f),
// where f marks a lane that begins a segment. Lanes without a source see
// (0, 0), which leaves them unchanged.
template<int CTRL, int ROW_MASK, typename V, int KMAX>
__device__ __forceinline__ void seg_scan_step(V (&a)[KMAX], int& f, int k) {
    const int uf = dpp_int<CTRL, ROW_MASK>(f);
    V up[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        up[j] = (j < k) ? dpp_val<CTRL, ROW_MASK>(a[j]) : V(0);
    if (!f) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            a[j] = up[j] + a[j];
        f = uf;
    }
}

template<typename V, int KMAX>
__device__ __forceinline__ void wave_seg_scan(V (&a)[KMAX], int& f, int k) {
    seg_scan_step<kDppRowShr1, 0xf>(a, f, k);
    seg_scan_step<kDppRowShr2, 0xf>(a, f, k);
    seg_scan_step<kDppRowShr4, 0xf>(a, f, k);
    seg_scan_step<kDppRowShr8, 0xf>(a, f, k);
    seg_scan_step<kDppRowBcast15, 0xa>(a, f, k);
    seg_scan_step<kDppRowBcast31, 0xc>(a, f, k);
}

// value of a wave-uniform lane, as a scalar
__device__ __forceinline__ int lane_value(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template<typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        v += __shfl_xor(v, d, 64);
    return v;
}

// Matrix values and column indices are read once per product. A plain
// product streams them with non-temporal loads, so that they do not evict the
// x entries that neighbouring rows gather again from L2 (measured: 10% less
// HBM read traffic on the 7-point operator). Products fused with CG's dot
// (NT = false) use ordinary loads when the solver's working set, the matrix
// and four vectors, fits the 256 MB Infinity Cache: the solver multiplies the
// same matrix every iteration, and it then stays there (128^3 Poisson CG:
// 20.4 ms, 22.8 ms with non-temporal loads). Single-vector solves whose
// working set does not fit stream it non-temporally like a plain product
// (32 uneven 2-D systems, 427 MB: 0.586 s against 0.611 s); see
// solve_streams_matrix().
template<bool NT, typename V>
__device__ __forceinline__ V ldg_stream(const V* p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// Per-tile metadata read at wave-uniform addresses goes through the constant
// address space, so that the loads are scalar (SMEM) loads that neither take
// vector registers nor queue behind the vector loads; these arrays are never
// written while a product runs. Tile flags are read as the aligned 32-bit word
// that holds them (flag arrays carry at least 4 bytes of padding).
template<typename T>
__device__ __forceinline__ T sload(const T* p) {
    return *(const __attribute__((address_space(4))) T*)(p);
}
__device__ __forceinline__ uint32_t sload_flag(const uint8_t* p) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint32_t w = sload(reinterpret_cast<const uint32_t*>(a & ~uintptr_t{3}));
    return (w >> ((a & 3u) * 8u)) & 0xffu;
}

// per-chunk flags written by the tile kernel
constexpr uint8_t kChunkStarts = 1; // the chunk's first element starts a row
constexpr uint8_t kChunkHead = 2;   // head slot holds the closing part of a row from earlier chunks
constexpr uint8_t kChunkTail = 4;   // tail slot holds the open part of a row that continues
constexpr uint8_t kChunkRowStart = 8; // a row starts somewhere inside the chunk

struct TileParams {
    const uint32_t* tile_ptr;
    const void* bit_flag;
    const uint8_t* tile_flags;
    const uint32_t* tdo_ptr;
    const uint32_t* tdo;
    const uint32_t* col_idx;
    const uint32_t* col_code; // column codes and their diagonals, or null
    const int32_t* col_dict;
    uint32_t cols;
    const void* vals;
    const void* x;
    void* y;
    void* head;
    void* tail;
    uint8_t* chunk_flags;
    double* dot_chunk;
    const int* done;
    const uint32_t* empty_rows;
    uint64_t nnz;
    uint32_t tile_begin, tile_end, row_base, last_tile;
    uint32_t tpw, nchunks;
    uint32_t empty_begin, empty_end; // empty rows written by this kernel
    int k;
    int lookahead;
    // Chunk table of a block-diagonal batch (null: chunks of tpw tiles from
    // tile_begin): chunk c covers tiles [chunk_tile[c], chunk_tile[c + 1]) of
    // system chunk_sys[c], so no chunk holds tiles of two systems. The
    // dictionary of a coded chunk is then col_dict + kMaxDiagonals *
    // chunk_sys[c] (every system keeps its own diagonals), and a dot product
    // leaves one partial per chunk in dot_chunk[c]. Each system owns whole
    // tiles, so no row crosses from one system into another.
    const uint32_t* chunk_tile;
    const uint32_t* chunk_sys;
    // the systems' stop flags (single-vector products with a chunk table; may
    // be null): the chunks of a system whose flag is set are skipped
    const int* sys_done;
    // x and y hold k columns: element (i, j) of x at i*rs + j*xcs, of y at
    // i*rs + j*ycs (row-major n x K blocks: rs = K, xcs = ycs = 1;
    // column-major: rs = 1, xcs and ycs the leading dimensions)
    uint64_t rs, xcs, ycs;
    // Ordered sweep (multi-RHS products; null: one chunk per wavefront): the
    // chunks are cut in kXcds contiguous ranges, and the wavefronts of blocks
    // b with b % kXcds == x take the chunks of range x in order from the
    // counter queue[x * kQueueStride] (its count of finished wavefronts at
    // queue[x * kQueueStride + 1]), so that the wavefronts resident on one
    // XCD work on neighbouring chunks at any time.
    uint32_t* queue;
};


// Multi-RHS products: right-hand sides whose gathers are in flight together
// in the fast path, and the tallest tile whose matrix column a lane keeps in
// registers.
template<typename V>
constexpr int smm_group() {
    return sizeof(V) == 8 ? 2 : 4;
}
constexpr int kSmmRegSigma = 16;

// One wavefront's work on chunk c: the product over the chunk's tiles, its
// carries and look-ahead, its share of the empty rows, and for CG the dot
// partial of the rows it writes. Returns that partial summed over the
// wavefront for single-vector products with DOT (multi-RHS products store
// theirs per chunk); the caller stores it. `stopped` (single-vector products
// with a chunk table): the chunk's system has stopped, so its rows are left
// as they are (a stopped system's product is not used); the chunk still
// writes its flags, as a chunk that starts a row and takes no carry, and its
// share of the empty rows. The fix-up kernel may then add stale head and
// tail slots of such a system's chunks into its rows of y; nothing reads
// them.
// CODES: the matrix carries column codes (every tile clean, sigma <= 8); the
// kernel then streams one code word per lane instead of sigma column indices.
template<typename V, int SIGMA, int KMAX, bool DOT, bool CODES = false, bool NT = !DOT>
__device__ __forceinline__ double csr5_chunk(const TileParams& p, const uint32_t c, const bool stopped) {
    const int lane = threadIdx.x & 63;
    uint32_t t0, t1;
    if (p.chunk_tile) {
        t0 = sload(p.chunk_tile + c);
        t1 = sload(p.chunk_tile + c + 1);
    } else {
        t0 = p.tile_begin + c * p.tpw;
        t1 = min(t0 + p.tpw, p.tile_end);
    }
    if constexpr (KMAX == 1) {
        if (stopped) {
            if (lane == 0)
                p.chunk_flags[c] = kChunkStarts | kChunkRowStart;
            if (p.lookahead) {
                V* __restrict__ yy = static_cast<V*>(p.y);
                for (uint32_t i = p.empty_begin + c * 64 + lane; i < p.empty_end; i += p.nchunks * 64)
                    yy[static_cast<size_t>(p.empty_rows[i] - p.row_base)] = V(0);
            }
            return 0.0; // wave-uniform
        }
    }

    using W = typename DescWord<SIGMA>::type;
    const int k = KMAX == 1 ? 1 : p.k;
    const V* __restrict__ x = static_cast<const V*>(p.x);
    V* __restrict__ y = static_cast<V*>(p.y);
    const size_t rs = KMAX == 1 ? 1 : p.rs;   // row stride of x and y
    const size_t xcs = KMAX == 1 ? 0 : p.xcs; // column strides
    const size_t ycs = KMAX == 1 ? 0 : p.ycs;

    // The row left open by the previous tile of this chunk (its partial sum),
    // and whether that row began before the chunk: its partial then goes to
    // the chunk's head slot instead of y.
    V carry[KMAX];
    double dacc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        carry[j] = V(0);
        dacc[j] = 0.0;
    }
    // Coded single-vector products with the dot (DIAGX below): x at the
    // diagonal element of the carried row, when an earlier tile held it.
    V carry_dx = V(0);
    bool carry_dx_ok = false;
    const uint32_t tf0 = sload_flag(p.tile_flags + t0);
    const bool chunk_starts = tf0 & kStartsRow;
    bool carry_is_head = !chunk_starts;
    uint8_t cflags = chunk_starts ? kChunkStarts : 0;
    bool row_start_seen = false;

    // Look-ahead (single-vector products): a row that starts in the last tile
    // of a chunk and ends in the first tile of the next chunk is finished by
    // the earlier chunk, which reads the few elements of that row in the next
    // tile; the later chunk drops its part. Other rows that cross chunks go
    // through the head/tail slots and the fix-up kernel. Both chunks decide
    // from the same two tile flags.
    constexpr bool LOOK = KMAX == 1;
    const bool look = LOOK && p.lookahead;
    bool drop_head = false;
    if (look && !chunk_starts && t0 > p.tile_begin && (tf0 & kMultiSeg))
        drop_head = (sload_flag(p.tile_flags + t0 - 1) & (kStartsRow | kMultiSeg)) != 0;

    // A lane keeps its whole column of the tile in registers so that all its
    // loads are in flight at once (multi-RHS products up to sigma 16; taller
    // tiles stream the column step by step). Single-vector products also
    // gather their sigma x entries up front (XREG) and, for short tiles, run
    // a two-stage pipeline over the chunk: the sigma values, column indices,
    // bit flags and tile metadata of tile t+1 are in flight while the x
    // gathers and the scans of tile t run, so the stream of matrix bytes does
    // not stop while a wavefront works on a tile (taller tiles would lose more
    // to the lower occupancy than they gain). Multi-RHS products gather per
    // right-hand side inside the tile, after the decode.
    constexpr bool XREG = KMAX == 1;
    constexpr bool REG = XREG || SIGMA <= kSmmRegSigma;
    // Plain single-vector products stage the row sums of a clean tile's
    // general path in LDS (one region per wavefront of the 256-thread block)
    // and write the tile's rows with one coalesced pass: written where they
    // are closed, the sums of short rows reach y as scattered 8-byte stores
    // that fill each line of y piecemeal. Measured on the MI355X, power-law
    // product (column-sliced, 2.5 elements per stacked row): those stores
    // took 135 of its 596 us.
    constexpr bool STAGE = KMAX == 1 && !DOT && SIGMA <= 8;
    V* ystage = nullptr;
    if constexpr (STAGE) {
        __shared__ V ybuf[4][64 * SIGMA];
        ystage = ybuf[(threadIdx.x >> 6) & 3];
    }
    constexpr bool PIPE = XREG && SIGMA <= 8;
    constexpr bool CODED = CODES && REG && SIGMA <= 8;
    // The dot needs x at every row the product writes. A coded single-vector
    // product takes it from its own gathers where it can: the element on the
    // diagonal (code naming offset 0) gathered x at its row. Measured on the
    // first product of a CG solve: loading x at the rows instead costs 42.6
    // against 38.7 us (128^3 Poisson), 81.7 against 73.9 us (32 uneven 2-D
    // systems). Keeping those x values takes 9 more VGPRs, which at tile
    // heights 4 and 7 costs a wavefront per SIMD (90 -> 99 and 125 -> 134
    // VGPRs); at 7 that made the 128^3 Poisson CG slower (20.8 -> 21.9 ms),
    // so only tile heights 5 and 6 (4 wavefronts per SIMD either way) take
    // it: 32 uneven 2-D systems at tile height 5, 0.603 -> 0.587 s per solve
    // (0.590 s with the diagonal mask formed after the gathers instead of
    // with the decode).
    constexpr bool DIAGX = DOT && KMAX == 1 && CODED && (SIGMA == 5 || SIGMA == 6);
    constexpr int NREG = REG ? SIGMA : 1;
    constexpr int NX = XREG ? SIGMA : 1;
    const V* __restrict__ vals_all = static_cast<const V*>(p.vals);
    const W* __restrict__ bits_all = static_cast<const W*>(p.bit_flag);
    // the diagonals of a coded matrix (of the chunk's system in a batch),
    // lane d holding diagonal d
    const int dtab =
      CODED && lane < 16 ? p.col_dict[(p.chunk_sys ? sload(p.chunk_sys + c) * kMaxDiagonals : 0u) + lane] : 0;

    struct Stage {
        V v[NREG];
        uint32_t col[NREG];
        uint32_t code;
        uint32_t dg; // DIAGX: steps whose element lies on the diagonal
        W bits;
        uint32_t tile_ptr, tdo_at;
        uint32_t tf, tf_next;
    };
    // Tile loads add a 32-bit lane offset to a wave-uniform tile base, so that
    // they take the base from scalar registers (the readfirstlane keeps the
    // compiler from turning the bases into per-lane 64-bit pointers).
    const uint32_t ulane = static_cast<uint32_t>(lane);
    auto fetch = [&](uint32_t t, Stage& s) {
        const size_t tile64 = static_cast<size_t>(__builtin_amdgcn_readfirstlane(t)) * 64;
        const size_t base = tile64 * SIGMA;
        if constexpr (REG) {
            const V* vt = vals_all + base;
            const uint32_t* ct = p.col_idx + base;
#pragma unroll
            for (int r = 0; r < SIGMA; ++r) {
                s.v[r] = ldg_stream<NT>(vt + r * 64 + ulane);
                if constexpr (!CODED)
                    s.col[r] = ldg_stream<NT>(ct + r * 64 + ulane);
            }
            if constexpr (CODED)
                s.code = ldg_stream<NT>(p.col_code + tile64 + ulane);
        }
        if constexpr (!CODED) // a code word carries its lane's flags
            s.bits = (bits_all + tile64)[ulane];
        // tile_flags and tile_desc_offset_ptr hold T + 1 entries
        s.tile_ptr = sload(p.tile_ptr + t);
        s.tdo_at = CODED ? 0u : sload(p.tdo_ptr + t); // coded matrices have no dirty tiles
        s.tf = sload_flag(p.tile_flags + t);
        s.tf_next = sload_flag(p.tile_flags + t + 1);
    };
    // Columns of a coded tile (the code word also carries the lane's row-start
    // flags): the element at step r of a lane belongs to row
    // row0 + y_off - 1 + (row starts of the lane at steps <= r), and its
    // column is that row plus the diagonal its code names. Clamped to the
    // column range, so that the padding of the last tile (code 0) reads a
    // valid x entry; padded elements contribute zero regardless.
    auto decode = [&](Stage& s) {
        if constexpr (CODED) {
            constexpr uint32_t CB = code_bits(SIGMA);
            s.bits = static_cast<W>(s.code >> kCodeFlagShift);
            const uint32_t bits = static_cast<uint32_t>(s.bits) | (lane == 0 ? 1u : 0u);
            const int nb = __popc(bits);
            const int y_off = wave_incl_scan(nb, lane) - nb;
            const uint32_t rb = (__builtin_amdgcn_readfirstlane(s.tile_ptr) & ~kDirty) + static_cast<uint32_t>(y_off) - 1u;
            // all table lookups in flight together, then the adds
            int d[SIGMA];
#pragma unroll
            for (int r = 0; r < SIGMA; ++r)
                d[r] = __builtin_amdgcn_ds_bpermute(static_cast<int>((s.code >> (CB * r)) & ((1u << CB) - 1u)) << 2,
                                                    dtab);
#pragma unroll
            for (int r = 0; r < SIGMA; ++r) {
                const uint32_t row = rb + static_cast<uint32_t>(__popc(bits & ((2u << r) - 1u)));
                s.col[r] = min(row + static_cast<uint32_t>(d[r]), p.cols - 1u);
            }
            // (with the decode, while the tile's gathers are in flight)
            if constexpr (DIAGX) {
                uint32_t dg = 0;
#pragma unroll
                for (int r = 0; r < SIGMA; ++r)
                    dg |= (d[r] == 0 ? 1u : 0u) << r;
                s.dg = dg;
            }

        }
    };
    // Up to sigma 7 the single-vector pipeline runs two tiles ahead (two
    // stage slots, the chunk walked in pairs so that each slot keeps its
    // registers). Measured on the 7-point operator: coded sigma 7 gains 4%,
    // coded sigma 8 loses 5% to the lower occupancy; 32-bit columns at
    // sigma 7: 45.8 -> 45.0 us.
    constexpr bool PIPE2 = PIPE && SIGMA <= 7;
    constexpr uint32_t AHEAD = PIPE2 ? 2 : 1;
    Stage nx, nx2;
    if constexpr (PIPE)
        fetch(t0, nx);

    // How the chunk ends follows from two tile flags: it ends inside a row
    // (end_open) unless its last tile's last element ends a row; with
    // look-ahead, the row is finished here when it started in the chunk's
    // last tile and ends inside tile t1 (end_look). The row-start flags of
    // tile t1, the row and, for the dot, its x entry are loaded now, so that
    // the look-ahead after the loop waits for fewer dependent loads.
    const uint32_t tf_last = sload_flag(p.tile_flags + t1 - 1);
    const uint32_t tf_end = sload_flag(p.tile_flags + t1); // tile_flags holds T + 1 entries
    const bool end_open = !((t1 == p.tile_end) || (tf_end & kStartsRow));
    const bool end_look =
      look && end_open && (tf_last & (kStartsRow | kMultiSeg)) != 0 && (tf_end & kMultiSeg) != 0;
    W wb_end = 0;
    uint32_t row_end = 0;
    V x_end = V(0);
    if constexpr (LOOK) {
        if (end_look) {
            wb_end = bits_all[static_cast<size_t>(t1) * 64 + lane];
            row_end = sload(p.tile_ptr + t1) & ~kDirty;
            if constexpr (DOT)
                x_end = x[row_end];
        }
    }
    // The elements of tile t1 before its first row start: all of the lanes
    // before the first lane with a row start, and that lane's elements before
    // its first flag. Set once the flags have arrived (with the first tile's
    // loads), before the tile loop.
    uint32_t m_end = 0;
    auto make_look_mask = [&]() {
        if constexpr (LOOK) {
            if (end_look) {
                const uint64_t any = __ballot(wb_end != 0);
                const int ls = __ffsll(static_cast<long long>(any)) - 1;
                const uint32_t bl = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(wb_end), ls));
                const int rs = __ffs(static_cast<int>(bl)) - 1;
                constexpr uint32_t full = static_cast<uint32_t>((uint64_t{1} << SIGMA) - 1u);
                m_end = lane < ls ? full : (lane == ls ? ((1u << rs) - 1u) : 0u);
            }
        }
    };
    // Pipelined kernels load those elements (values, and column indices or
    // the lane's code word) into the stage slot that the step of tile
    // t1 - AHEAD would have prefetched into, so that after the loop only
    // their x gathers remain.
    auto fetch_look = [&](Stage& s) {
        if constexpr (PIPE) {
            const size_t tile64 = static_cast<size_t>(__builtin_amdgcn_readfirstlane(t1)) * 64;
            const size_t b1 = tile64 * SIGMA;
#pragma unroll
            for (int r = 0; r < SIGMA; ++r) {
                if ((m_end >> r) & 1u) {
                    s.v[r] = vals_all[b1 + r * 64 + ulane];
                    if constexpr (!CODED)
                        s.col[r] = p.col_idx[b1 + r * 64 + ulane];
                }
            }
            if constexpr (CODED) {
                if (m_end)
                    s.code = p.col_code[tile64 + ulane];
            }
        }
    };

    // pf_tag true: the tile t + AHEAD exists, so the prefetch is issued on
    // every path (the waits for a slot's loads then have a fixed count of
    // later loads to leave in flight)
    auto tile_step = [&](uint32_t t, Stage& slot, auto pf_tag) {
        constexpr bool IN_RANGE = decltype(pf_tag)::value;
        const size_t base = static_cast<size_t>(t) * (64 * SIGMA);
        const V* vals = vals_all + base + lane;
        const uint32_t* cols = p.col_idx + base + lane;

        Stage cu;
        V xg[NX];
        if constexpr (PIPE) {
            cu = slot;
            decode(cu);
#pragma unroll
            for (int r = 0; r < SIGMA; ++r) {
                if constexpr (CODED) // x of a coded matrix spans less than 4 GB: scalar base + 32-bit offset
                    xg[r] = *reinterpret_cast<const V*>(reinterpret_cast<const char*>(x) +
                                                        cu.col[r] * static_cast<uint32_t>(sizeof(V)));
                else
                    xg[r] = x[cu.col[r]];
            }
            // keep the next tiles' loads behind this tile's gathers: loads
            // return in order, so waiting for the gathers then leaves the next
            // tiles' loads in flight
            __builtin_amdgcn_sched_barrier(0);
            if (IN_RANGE || t + AHEAD < t1)
                fetch(t + AHEAD, slot);
            else if (LOOK && end_look && t + AHEAD == t1)
                fetch_look(slot);
        } else {
            fetch(t, cu);
            decode(cu);
            if constexpr (XREG) {
#pragma unroll
                for (int r = 0; r < SIGMA; ++r)
                    xg[r] = x[cu.col[r]];
            }
        }

        const uint32_t bits = static_cast<uint32_t>(cu.bits) | (lane == 0 ? 1u : 0u);
        const uint32_t tf = __builtin_amdgcn_readfirstlane(cu.tf);
        const bool starts_row = tf & kStartsRow;
        const uint32_t tpw = __builtin_amdgcn_readfirstlane(cu.tile_ptr);
        const bool dirty = (tpw & kDirty) != 0;
        const uint32_t row0 = tpw & ~kDirty;
        const uint32_t tf_next = __builtin_amdgcn_readfirstlane(cu.tf_next);
        const bool complete = (t + 1 == p.tile_end) || (tf_next & kStartsRow);
        const bool has_row_start = (tf & (kStartsRow | kMultiSeg)) != 0;
        row_start_seen |= has_row_start;

        // padded elements of the last tile contribute exactly zero
        uint32_t pad = 0;
        if (t == p.last_tile) {
            const uint32_t left = static_cast<uint32_t>(p.nnz - base); // elements of the tile that exist
#pragma unroll
            for (int r = 0; r < SIGMA; ++r)
                if (ulane * SIGMA + r >= left) {
                    pad |= 1u << r;
                    if constexpr (XREG)
                        xg[r] = V(0);
                }
        }

        const int nb = __popc(bits);

        // Fast path for the common tile of short rows: no empty rows, no
        // padding, and one or two row starts in every lane. The step loop then
        // needs no branches: the lane's leading partial (red) closes the row
        // left open by the lane below, a second start closes one row inside
        // the lane (mid), and the trailing partial continues into the lane
        // above. Multi-RHS products decode the tile once and run the step loop
        // once per right-hand side, G sides at a time with their gathers in
        // flight together.
        if (REG && !dirty && t != p.last_tile && __ballot(nb == 0 || nb > 2) == 0ull) {
            if constexpr (REG) {
                const int f1 = __builtin_ctz(bits);
                const uint32_t rest = bits & (bits - 1u);
                const int f2 = rest ? __builtin_ctz(rest) : 64;
                const int incl = wave_incl_scan(nb, lane);
                const int y_off = incl - nb; // segment of the lane's first start
                const int nseg = lane_value(incl, 63);
                const size_t yrow0 = static_cast<size_t>(row0 - p.row_base);
                // segment 0 is lane 0's mid when lane 0 holds two starts,
                // otherwise lane 1's tot; it continues the carried row unless
                // the tile starts a row
                const bool two0 = lane_value(nb, 0) == 2;
                const bool keep0 = starts_row;
                // The rows a lane may write: segment y_off - 1 (tot), y_off
                // (mid; in lane 0 segment 0; in lane 63 with one start the
                // tile's last segment) and y_off + 1 (the tile's last segment
                // in lane 63 with two starts). The fused dot needs x at them:
                // a lane loads x at its own starts (rows y_off and, with two
                // starts, y_off + 1) and takes x at row y_off - 1, the last
                // start of the lane below (every lane holds one or two
                // starts here), from that lane.
                const size_t rb = row0 + static_cast<uint32_t>(y_off);
                // DIAGX: the lane's diagonal elements among its head (steps
                // before its first start: the row of the lane below), its first
                // row (complete in the lane when it holds two starts) and its
                // last row. The row a lane closes for the lane below has its
                // diagonal in the head or in the lane below's tail; segment 0
                // of a tile that continues a row, in an earlier tile, in lane
                // 0, or (lane 0 holding that row alone) in lane 1's head; the
                // tile's last row, in lane 63's tail. A wave where some written
                // row has none of these (a row without a stored diagonal)
                // loads x at the rows instead.
                uint32_t hm = 0, fm = 0, tm = 0;
                bool dfall = false;
                int fv0 = 0, hv1 = 0;
                if constexpr (DIAGX) {
                    const uint32_t dg = cu.dg;
                    const uint32_t below1 = (1u << f1) - 1u;
                    hm = dg & below1;
                    fm = dg & ~below1 & (nb == 2 ? ((1u << f2) - 1u) : ~0u);
                    tm = dg & ~((1u << (nb == 2 ? f2 : f1)) - 1u);
                    const int tv_below = dpp_int<kDppWaveShr1, 0xf>(tm != 0 ? 1 : 0);
                    fv0 = lane_value(fm != 0 ? 1 : 0, 0);
                    hv1 = lane_value(hm != 0 ? 1 : 0, 1);
                    const bool need = (lane >= 1 && (keep0 || lane != 1 || two0) && hm == 0 && !tv_below) ||
                                      (nb == 2 && (keep0 || lane != 0) && fm == 0) ||
                                      (lane == 0 && !keep0 && !carry_is_head && !carry_dx_ok && !fv0 &&
                                       (two0 || !hv1)) ||
                                      (lane == 63 && complete && tm == 0);
                    dfall = __ballot(need) != 0ull;
                }
                V td_last = V(0); // DIAGX: x at the diagonal of the lane's last row (side 0)
                constexpr int G = KMAX < smm_group<V>() ? KMAX : smm_group<V>();
#pragma unroll
                for (int g = 0; g < KMAX; g += G) {
                    if (g < k) {
                        // sides past k re-read side k - 1 and store nothing
                        V xv[G][SIGMA];
                        V px[G][2]; // x at the lane's first and last row starts (the dot)
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const V* xj = x + static_cast<size_t>(g + i < k ? g + i : k - 1) * xcs;
#pragma unroll
                            for (int r = 0; r < SIGMA; ++r) {
                                if constexpr (XREG)
                                    xv[i][r] = xg[r];
                                else
                                    xv[i][r] = xj[static_cast<size_t>(cu.col[r]) * rs];
                            }
                            if constexpr (DOT && !DIAGX) {
                                // no branch between the gathers: a lane with
                                // one start loads row y_off twice
                                px[i][0] = xj[rb * rs];
                                px[i][1] = xj[(rb + static_cast<uint32_t>(nb - 1)) * rs];
                            }
                        }
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const int j = g + i;
                            if (j < k) {
                                V acc = V(0), red = V(0), mid = V(0);
#pragma unroll
                                for (int r = 0; r < SIGMA; ++r) {
                                    acc = fma(cu.v[r], xv[i][r], (r == f1 || r == f2) ? V(0) : acc);
                                    red = (r + 1 == f1) ? acc : red;
                                    mid = (r + 1 == f2) ? acc : mid;
                                }
                                const V tot = dpp_val<kDppWaveShr1, 0xf>(acc) + red; // lanes >= 1: segment y_off - 1
                                // x at the rows this lane writes: row y_off - 1
                                // (pa), its first start's row (pf), its last
                                // start's row (pl) and, in lane 0, segment 0's
                                // row (ps0). Without DIAGX, x at row y_off - 1
                                // is the lane below's last-start load, shifted
                                // here, after the gathers have been waited for,
                                // not between them (FP64 k = 8 with the dot:
                                // 580 us, against 695 us with the shift right
                                // after the loads and 618 us with three
                                // gathers per lane).
                                V pa = V(0), pf = V(0), pl = V(0), ps0 = V(0);
                                if constexpr (DIAGX) {
                                    V hd = V(0), fd = V(0), td = V(0); // x at the diagonal elements
#pragma unroll
                                    for (int r = 0; r < SIGMA; ++r) {
                                        hd = ((hm >> r) & 1u) ? xv[i][r] : hd;
                                        fd = ((fm >> r) & 1u) ? xv[i][r] : fd;
                                        td = ((tm >> r) & 1u) ? xv[i][r] : td;
                                    }
                                    const V td_below = dpp_val<kDppWaveShr1, 0xf>(td);
                                    pa = hm ? hd : td_below;
                                    pf = fd;
                                    pl = td;
                                    const V fd0 = lane_value(fd, 0), hd1 = lane_value(hd, 1);
                                    ps0 = carry_dx_ok ? carry_dx : (fv0 ? fd0 : hd1);
                                    td_last = td;
                                    if (dfall) { // wave-uniform
                                        pf = ps0 = x[rb];
                                        pl = x[rb + static_cast<uint32_t>(nb - 1)];
                                        pa = dpp_val<kDppWaveShr1, 0xf>(pl);
                                    }
                                } else if constexpr (DOT) {
                                    pa = dpp_val<kDppWaveShr1, 0xf>(px[i][1]);
                                    pf = ps0 = px[i][0];
                                    pl = px[i][1];
                                }
                                V* yj = y + static_cast<size_t>(j) * ycs;
                                auto store_row = [&](int seg, V val, V xr) {
                                    yj[(yrow0 + seg) * rs] = val;
                                    if constexpr (DOT)
                                        dacc[j] += static_cast<double>(xr) * static_cast<double>(val);
                                };
                                if (lane >= 1 && (keep0 || lane != 1 || two0))
                                    store_row(y_off - 1, tot, pa);
                                if (nb == 2 && (keep0 || lane != 0))
                                    store_row(y_off, mid, pf);
                                if (!keep0) {
                                    const V s0 = carry[j] + (two0 ? lane_value(mid, 0) : lane_value(tot, 1));
                                    if (lane == 0) {
                                        if (!carry_is_head)
                                            store_row(0, s0, ps0);
                                        else if (!drop_head)
                                            static_cast<V*>(p.head)[static_cast<size_t>(c) * k + j] = s0;
                                    }
                                }
                                // the tile's last segment starts in lane 63
                                const V tv = lane_value(acc, 63);
                                if (complete) {
                                    if (lane == 63)
                                        store_row(nseg - 1, tv, pl);
                                    carry[j] = V(0);
                                } else {
                                    carry[j] = tv;
                                }
                            }
                        }
                    }
                }
                if (!keep0)
                    carry_is_head = false;
                if constexpr (DIAGX) {
                    // the row carried into the next tile started in lane 63
                    carry_dx_ok = !complete && lane_value(tm != 0 ? 1 : 0, 63) != 0;
                    carry_dx = lane_value(td_last, 63);
                }
            }
        } else {
        const int incl = wave_incl_scan(nb, lane);
        const int y_off = incl - nb; // first segment that starts in this lane
        const int nseg = lane_value(incl, 63);
        const uint64_t started = __ballot(nb > 0);

        // The rest of the tile is instantiated twice: tiles without empty rows
        // map segment s to row row0 + s with no memory access, so no load (and
        // no wait for loads) sits between the gathers and the stores.
        // Multi-RHS tiles run it once per group of GG right-hand sides, so that
        // only GG partial sums of each kind are live in a lane.
        auto process = [&](auto dirty_tag) {
            constexpr bool DIRTY = decltype(dirty_tag)::value;
            constexpr int GG = KMAX < smm_group<V>() ? KMAX : smm_group<V>();
            struct Vec {
                V a[GG];
            };
            const uint32_t* tdo = DIRTY ? p.tdo + __builtin_amdgcn_readfirstlane(cu.tdo_at) : nullptr;
            auto row_of = [&](int s) -> uint32_t {
                if constexpr (DIRTY)
                    return row0 + tdo[s];
                else
                    return row0 + static_cast<uint32_t>(s);
            };
            const bool cih = carry_is_head;
#pragma unroll
            for (int g = 0; g < KMAX; g += GG) {
                if (g >= k)
                    break; // wave-uniform
                const int kg = k - g; // sides i < kg of this group exist
                auto write_row = [&](uint32_t row, const Vec& val) {
                    V* yr = y + static_cast<size_t>(row - p.row_base) * rs + static_cast<size_t>(g) * ycs;
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        if (i < kg)
                            yr[i * ycs] = val.a[i];
                    if constexpr (DOT) {
                        const V* pr = x + static_cast<size_t>(row) * rs + static_cast<size_t>(g) * xcs;
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            if (i < kg)
                                dacc[g + i] += static_cast<double>(pr[i * xcs]) * static_cast<double>(val.a[i]);
                    }
                };
                // segment 0 of a tile that does not start a row closes the carried row
                auto close_carried = [&](const Vec& val) {
                    Vec tot;
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        tot.a[i] = carry[g + i] + val.a[i];
                    if (cih) {
                        if (drop_head)
                            return; // the previous chunk writes this row
                        V* h = static_cast<V*>(p.head) + static_cast<size_t>(c) * k + g;
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            if (i < kg)
                                h[i] = tot.a[i];
                    } else {
                        write_row(row0, tot);
                    }
                };
                // Segment 0 of a tile that does not start a row continues the
                // carried row. The lane that closes it only keeps its partial;
                // one lane adds it to the carry after the scan, so that the head
                // slot and y are written from wave-uniform code only.
                bool own0 = false;
                Vec seg0;
#pragma unroll
                for (int i = 0; i < GG; ++i)
                    seg0.a[i] = V(0);
                auto emit = [&](int s, const Vec& val) {
                    if (s == 0 && !starts_row) {
                        own0 = true;
                        seg0 = val;
                    } else if constexpr (STAGE && !DIRTY) {
                        ystage[s] = val.a[0];
                    } else {
                        write_row(row_of(s), val);
                    }
                };

                Vec cur, red;
#pragma unroll
                for (int i = 0; i < GG; ++i) {
                    cur.a[i] = V(0);
                    red.a[i] = V(0);
                }
                bool has = false;
                int seg = y_off - 1;
                auto step = [&](int r, V vr, uint32_t cr) {
                    if ((bits >> r) & 1u) {
                        if (has)
                            emit(seg, cur);
                        else
                            red = cur;
                        has = true;
                        ++seg;
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            cur.a[i] = V(0);
                    }
                    if constexpr (XREG) {
                        cur.a[0] = fma(vr, xg[r], cur.a[0]); // padded steps hold x = 0
                    } else if (!((pad >> r) & 1u)) {
                        const V* xr = x + static_cast<size_t>(cr) * rs + static_cast<size_t>(g) * xcs;
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            if (i < kg)
                                cur.a[i] = fma(vr, xr[i * xcs], cur.a[i]);
                    }
                };
                if constexpr (REG) {
#pragma unroll
                    for (int r = 0; r < SIGMA; ++r)
                        step(r, cu.v[r], XREG ? 0u : cu.col[r]);
                } else {
#pragma unroll 1
                    for (int r = 0; r < SIGMA; ++r)
                        step(r, ldg_stream<NT>(vals + r * 64), ldg_stream<NT>(cols + r * 64));
                }

                V blue[GG];
#pragma unroll
                for (int i = 0; i < GG; ++i) {
                    if (has) {
                        blue[i] = cur.a[i];
                    } else {
                        red.a[i] = cur.a[i];
                        blue[i] = V(0);
                    }
                }

                // A lane in which a row starts closes the row left open by the
                // previous such lane: that lane's trailing (blue) partial plus
                // the leading (red) partials of the lanes in between, summed by
                // a segmented inclusive scan (a new segment begins after every
                // lane in which a row starts). When every lane starts a row
                // (rows no longer than sigma, the common case) the scan is the
                // identity and the previous such lane is the lane below: one
                // DPP shift.
                const bool all_lanes = started == ~0ull;
                V A[GG], bc[GG];
#pragma unroll
                for (int i = 0; i < GG; ++i)
                    A[i] = red.a[i];
                if (all_lanes) {
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        bc[i] = (i < kg) ? dpp_val<kDppWaveShr1, 0xf>(blue[i]) : V(0);
                } else {
                    int f = (lane == 0) ? 1 : static_cast<int>((started >> (lane - 1)) & 1ull);
                    wave_seg_scan<V, GG>(A, f, kg);
                    const uint64_t below = lane ? (started & ((1ull << lane) - 1ull)) : 0ull;
                    const int cprev = below ? 63 - __clzll(below) : 0;
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        bc[i] = __shfl(blue[i], cprev, 64);
                }
                if (lane > 0 && has) {
                    Vec tot;
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        tot.a[i] = bc[i] + A[i];
                    emit(y_off - 1, tot);
                }

                if (!starts_row && nseg > 1) {
                    // exactly one lane closed segment 0
                    const int o = __ffsll(static_cast<long long>(__ballot(own0))) - 1;
                    Vec s0;
#pragma unroll
                    for (int i = 0; i < GG; ++i)
                        s0.a[i] = (i < kg) ? lane_value(seg0.a[i], o) : V(0);
                    if (lane == 0)
                        close_carried(s0);
                }

                // The tile's last segment, known to every lane.
                const int L = all_lanes ? 63 : 63 - __clzll(started);
                Vec tv;
#pragma unroll
                for (int i = 0; i < GG; ++i) {
                    if (i < kg) {
                        const V bl = lane_value(blue[i], L);
                        const V a63 = lane_value(A[i], 63);
                        tv.a[i] = (L == 63) ? bl : bl + a63;
                    } else {
                        tv.a[i] = V(0);
                    }
                }
                if (nseg == 1 && !starts_row) {
                    // the whole tile continues the carried row
                    if (complete) {
                        if (lane == 0)
                            close_carried(tv);
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            carry[g + i] = V(0);
                    } else {
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            carry[g + i] = carry[g + i] + tv.a[i];
                    }
                } else {
                    if (complete) {
                        if (lane == 63)
                            emit(nseg - 1, tv);
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            carry[g + i] = V(0);
                    } else {
#pragma unroll
                        for (int i = 0; i < GG; ++i)
                            carry[g + i] = tv.a[i];
                    }
                }
                if constexpr (STAGE && !DIRTY) {
                    // Every segment in [s_lo, s_hi] was emitted exactly once
                    // (segment 0 of a tile that continues a row went to the
                    // carry or the head slot instead; the last segment stays
                    // open unless the tile completes it): write them in order.
                    // The wave barriers keep the compiler from moving these
                    // LDS accesses across the other lanes' stores and the next
                    // tile's; a wavefront's LDS instructions run in order.
                    __builtin_amdgcn_wave_barrier();
                    const int s_lo = starts_row ? 0 : 1;
                    const int s_hi = complete ? nseg - 1 : nseg - 2;
                    V* yrow = y + static_cast<size_t>(row0 - p.row_base);
                    for (int i = s_lo + lane; i <= s_hi; i += 64)
                        yrow[i] = ystage[i];
                    __builtin_amdgcn_wave_barrier();
                }
            }
            // segment 0 closed the carried row, or the whole tile did
            if (!starts_row && (nseg != 1 || complete))
                carry_is_head = false;
        };
        if (!CODED && dirty) // coded matrices have no dirty tiles
            process(std::true_type{});
        else
            process(std::false_type{});
        if constexpr (DIAGX)
            carry_dx_ok = false; // this path loads x at the rows it writes
        }

    };
    if constexpr (PIPE2) {
        // the main loop is entered with both slots loading on every path
        uint32_t t = t0;
        if (t0 + 3 < t1) {
            fetch(t0 + 1, nx2);
            make_look_mask();
            for (; t + 3 < t1; t += 2) {
                tile_step(t, nx, std::true_type{});
                tile_step(t + 1, nx2, std::true_type{});
            }
        } else {
            if (t0 + 1 < t1)
                fetch(t0 + 1, nx2);
            make_look_mask();
        }
        for (; t < t1; t += 2) {
            tile_step(t, nx, std::false_type{});
            if (t + 1 < t1)
                tile_step(t + 1, nx2, std::false_type{});
        }
    } else {
        make_look_mask();
        for (uint32_t t = t0; t < t1; ++t)
            tile_step(t, nx, std::false_type{});
    }
    if (end_open) {
        // the chunk ends inside a row
        if (end_look) {
            // the row started in the chunk's last tile and ends inside tile
            // t1: add the elements of tile t1 before its first row start
            if constexpr (LOOK) {
                V part = V(0);
                auto look_sum = [&](const Stage& st) {
                    // every element here lies in row row_end: a coded
                    // element's column is that row plus its diagonal
                    uint32_t cl[SIGMA];
#pragma unroll
                    for (int r = 0; r < SIGMA; ++r) {
                        if constexpr (CODED) {
                            constexpr uint32_t CB = code_bits(SIGMA);
                            const int d = __builtin_amdgcn_ds_bpermute(
                              static_cast<int>((st.code >> (CB * r)) & ((1u << CB) - 1u)) << 2, dtab);
                            cl[r] = min(row_end + static_cast<uint32_t>(d), p.cols - 1u);
                        } else {
                            cl[r] = st.col[r];
                        }
                    }
                    if (m_end) {
#pragma unroll
                        for (int r = 0; r < SIGMA; ++r)
                            if ((m_end >> r) & 1u)
                                part = fma(st.v[r], x[cl[r]], part);
                    }
                };
                // the step of tile t1 - AHEAD loaded them when it exists
                // (slot nx at an even offset from t0, nx2 at an odd one)
                if (PIPE && t1 >= t0 + AHEAD) {
                    if (PIPE2 && ((t1 - AHEAD - t0) & 1u))
                        look_sum(nx2);
                    else
                        look_sum(nx);
                } else if (m_end) {
                    const size_t b1 = static_cast<size_t>(t1) * (64 * SIGMA) + lane;
#pragma unroll
                    for (int r = 0; r < SIGMA; ++r)
                        if ((m_end >> r) & 1u)
                            part = fma(vals_all[b1 + r * 64], x[p.col_idx[b1 + r * 64]], part);
                }
                part = wave_sum(part);
                if (lane == 0) {
                    const V val = carry[0] + part;
                    y[static_cast<size_t>(row_end - p.row_base)] = val;
                    if constexpr (DOT)
                        dacc[0] += static_cast<double>(x_end) * static_cast<double>(val);
                }
            }
        } else {
            cflags |= kChunkTail;
            if (lane == 0) {
                V* tl = static_cast<V*>(p.tail) + static_cast<size_t>(c) * k;
#pragma unroll
                for (int j = 0; j < KMAX; ++j)
                    if (j < k)
                        tl[j] = carry[j];
            }
        }
    }
    if (!chunk_starts && !carry_is_head && !drop_head)
        cflags |= kChunkHead;
    if (row_start_seen)
        cflags |= kChunkRowStart;
    if (lane == 0)
        p.chunk_flags[c] = cflags;
    if (look) {
        // empty rows, spread over the chunks
        for (uint32_t i = p.empty_begin + c * 64 + lane; i < p.empty_end; i += p.nchunks * 64)
            y[static_cast<size_t>(p.empty_rows[i] - p.row_base)] = V(0);
    }
    if constexpr (DOT && KMAX == 1) {
        return wave_sum(dacc[0]);
    } else if constexpr (DOT) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const double s = wave_sum(dacc[j]);
            if (lane == 0 && j < k)
                p.dot_chunk[static_cast<size_t>(c) * k + j] = s;
        }
    }
    return 0.0;
}

template<typename V, int SIGMA, int KMAX, bool DOT, bool CODES = false, bool NT = !DOT>
__global__ __launch_bounds__(256) void k_csr5_tiles(TileParams p) {
    if (p.done && *p.done)
        return;
    if constexpr (KMAX > 1) {
        // One chunk per wavefront, or (queue set) the ordered sweep. Blocks
        // are dispatched to the XCDs in turn (block b on XCD b % kXcds): in
        // the sweep the wavefronts of one XCD share one range of chunks and
        // its counter, taking the chunks in order, so that their x gathers
        // and the rows they write stay within a window its L2 holds. Every
        // chunk is taken once; the wavefront that finishes last clears the
        // pair. Both forms share one loop, so the chunk code is emitted once.
        // Every lane runs the counter's atomic (lane 0 adds 1, the others 0),
        // so that the chunk index never comes out of a lane-0-only branch: a
        // loop whose exit followed a value produced under such a branch was
        // compiled with a per-lane exit that lanes 1..63 never took.
        const int lane = threadIdx.x & 63;
        const bool sweep = p.queue != nullptr;
        const uint32_t xr = blockIdx.x % kXcds;
        uint32_t* q = sweep ? p.queue + xr * kQueueStride : nullptr;
        // chunk c0 + i for i < n (the index is compared before it is added,
        // so no counter value can wrap into the range)
        uint32_t c0 = 0, n = p.nchunks, i;
        if (sweep) {
            const uint32_t per = (p.nchunks + kXcds - 1) / kXcds;
            c0 = min(xr * per, p.nchunks);
            n = min(c0 + per, p.nchunks) - c0;
            i = __builtin_amdgcn_readfirstlane(atomicAdd(q, lane == 0 ? 1u : 0u));
        } else {
            i = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
        }
        while (i < n) { // wave-uniform
            (void)csr5_chunk<V, SIGMA, KMAX, DOT, CODES, NT>(p, c0 + i, false);
            if (!sweep)
                break;
            i = __builtin_amdgcn_readfirstlane(atomicAdd(q, lane == 0 ? 1u : 0u));
        }
        if (sweep) {
            // wavefronts of this range: 4 per block b < gridDim.x with
            // b % kXcds == xr
            const uint32_t waves = 4u * (gridDim.x / kXcds + (xr < gridDim.x % kXcds ? 1u : 0u));
            if (lane == 0 && atomicAdd(q + 1, 1u) + 1u == waves) {
                atomicExch(q, 0u);
                atomicExch(q + 1, 0u);
            }
        }
        return;
    }
    // the chunk index is uniform across the wavefront: keep it (and every
    // per-tile quantity derived from it) in scalar registers
    const uint32_t c = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    if (c >= p.nchunks)
        return; // wave-uniform
    bool stopped = false;
    if constexpr (KMAX == 1) {
        if (p.chunk_sys && p.sys_done)
            stopped = sload(p.sys_done + sload(p.chunk_sys + c)) != 0;
    }
    const double s = csr5_chunk<V, SIGMA, KMAX, DOT, CODES, NT>(p, c, stopped);
    if constexpr (DOT && KMAX == 1) {
        const int lane = threadIdx.x & 63;
        if (p.chunk_tile) {
            // one partial per chunk: a batch sums them per system
            if (lane == 0)
                p.dot_chunk[c] = s;
            return;
        }
        // One partial per block of four chunks (block b: chunks 4b..4b+3, in
        // order), one per chunk in a last block that holds fewer than four:
        // dot_partials() of them.
        const uint32_t base = c & ~3u;
        if (base + 4 <= p.nchunks) { // block-uniform: every wave of the block is here
            __shared__ double red[4];
            if (lane == 0)
                red[c & 3u] = s;
            __syncthreads();
            if (threadIdx.x == 0)
                p.dot_chunk[base >> 2] = (red[0] + red[1]) + (red[2] + red[3]);
        } else if (lane == 0) {
            p.dot_chunk[(base >> 2) + (c & 3u)] = s;
        }
    }
}

// Dot partials the tile kernel leaves in dot_chunk: one per chunk for blocks
// of right-hand sides, one per block of four chunks (and per chunk of a last,
// partial block) for single vectors.
__host__ __device__ inline uint32_t dot_partials(int kmax, uint32_t nchunks) {
    if (kmax != 1)
        return nchunks;
    const uint32_t full = nchunks / 4, rest = nchunks % 4;
    return full + rest;
}

struct FixParams {
    const uint32_t* tile_ptr;
    const uint32_t* empty_rows;
    const uint8_t* chunk_flags;
    const void* x;
    void* y;
    const void* head;
    const void* tail;
    const double* dot_chunk;
    double* dot_blk;
    // chunk-table products with a single-vector dot: the dot partials stay
    // one per chunk, and the chunk that finishes a row adds that row's term
    // to its own partial here (null: partials are reduced into dot_blk)
    double* dot_rw;
    const int* done;
    const uint32_t* chunk_tile; // as TileParams::chunk_tile
    uint32_t tile_begin, row_base, empty_begin, empty_end;
    uint32_t tpw, nchunks;
    uint32_t ndot;               // dot partials in dot_chunk
    int k;
    uint64_t rs, xcs, ycs;       // as TileParams::rs, xcs, ycs
    uint32_t dot_ld, dot_off;    // dot partial of column j goes to dot_blk[block*dot_ld + dot_off + j]
};

// Rows that cross chunk boundaries: the chunk in which such a row ends adds
// the chain tail(start chunk) + tail(chunks inside the row) + head(end chunk)
// in chunk order and writes the row once. Also writes 0 to empty rows and,
// for CG, reduces the per-chunk dot partials to one partial per block.
template<typename V, int KMAX, bool DOT>
__global__ __launch_bounds__(256) void k_csr5_fixup(FixParams p) {
    if (p.done && *p.done)
        return;
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    const int k = KMAX == 1 ? 1 : p.k;
    const V* x = static_cast<const V*>(p.x);
    V* y = static_cast<V*>(p.y);
    const size_t rs = KMAX == 1 ? 1 : p.rs;
    const size_t xcs = KMAX == 1 ? 0 : p.xcs;
    const size_t ycs = KMAX == 1 ? 0 : p.ycs;
    const V* head = static_cast<const V*>(p.head);
    const V* tail = static_cast<const V*>(p.tail);
    const bool per_chunk = DOT && KMAX == 1 && p.dot_rw; // block-uniform
    double dacc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        dacc[j] = 0.0;

    if (c < p.nchunks) {
        const uint8_t cf = p.chunk_flags[c];
        if (cf & kChunkHead) {
            // walk back over chunks that lie entirely inside this row
            int64_t kk = static_cast<int64_t>(c) - 1;
            while (kk >= 0 && !(p.chunk_flags[kk] & kChunkRowStart))
                --kk;
            const uint32_t from = kk >= 0 ? static_cast<uint32_t>(kk) : 0u;
            const uint32_t row = p.tile_ptr[p.chunk_tile ? p.chunk_tile[c] : p.tile_begin + c * p.tpw] & ~kDirty;
            V* yr = y + static_cast<size_t>(row - p.row_base) * rs;
#pragma unroll
            for (int j = 0; j < KMAX; ++j) {
                if (j < k) {
                    V s;
                    if (from < c) {
                        s = tail[static_cast<size_t>(from) * k + j];
                        for (uint32_t u = from + 1; u < c; ++u)
                            s = s + tail[static_cast<size_t>(u) * k + j];
                        s = s + head[static_cast<size_t>(c) * k + j];
                    } else {
                        s = head[static_cast<size_t>(c) * k + j];
                    }
                    yr[j * ycs] = s;
                    if constexpr (DOT) {
                        const double term =
                          static_cast<double>(x[static_cast<size_t>(row) * rs + j * xcs]) * static_cast<double>(s);
                        if (per_chunk)
                            p.dot_rw[c] += term;
                        else
                            dacc[j] += term;
                    }
                }
            }
        }
    }
    if constexpr (DOT) {
        if (!per_chunk && c < p.ndot) {
#pragma unroll
            for (int j = 0; j < KMAX; ++j)
                if (j < k)
                    dacc[j] += p.dot_chunk[static_cast<size_t>(c) * k + j];
        }
    }

    const uint32_t ne = p.empty_end - p.empty_begin;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ne; i += gridDim.x * 256) {
        const uint32_t row = p.empty_rows[p.empty_begin + i];
        V* yr = y + static_cast<size_t>(row - p.row_base) * rs;
        for (int j = 0; j < k; ++j)
            yr[j * ycs] = V(0);
    }

    if constexpr (DOT) {
        if (per_chunk)
            return;
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
            p.dot_blk[static_cast<size_t>(blockIdx.x) * p.dot_ld + p.dot_off + threadIdx.x] =
              (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
    }
}

// Launch geometry: one chunk per wavefront, as many wavefronts as the device
// keeps resident for this kernel.
struct Plan {
    uint32_t tpw = 1, nchunks = 0, tgrid = 0, fgrid = 1;
    bool sweep = false; // multi-RHS ordered sweep (TileParams::queue)
};

int device_cus();

template<typename V, int SIGMA, int KMAX, bool DOT, bool CODES>
int tile_blocks_per_cu() {
    static int blocks_per_cu = 0;
    if (!blocks_per_cu) {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_csr5_tiles<V, SIGMA, KMAX, DOT, CODES>, 256, 0) !=
              hipSuccess ||
            b <= 0)
            b = 2;
        blocks_per_cu = b;
    }
    return blocks_per_cu;
}

// Tiles per chunk of the multi-RHS products' ordered sweep.
uint32_t sweep_tiles();

template<typename V, int SIGMA, int KMAX, bool DOT, bool CODES>
Plan make_plan(uint32_t ntiles) {
    const int blocks_per_cu = tile_blocks_per_cu<V, SIGMA, KMAX, DOT, CODES>();
    Plan pl;
    if (!ntiles)
        return pl;
    const uint64_t waves = static_cast<uint64_t>(device_cus()) * blocks_per_cu * 4;
    if (KMAX > 2 && sweep_tiles()) {
        // ordered sweep: short chunks, every resident wavefront looping
        // (from 3 right-hand sides: at 2 the gathers of one chunk per
        // wavefront already stay in L2)
        pl.sweep = true;
        pl.tpw = sweep_tiles();
        pl.nchunks = (ntiles + pl.tpw - 1) / pl.tpw;
        pl.tgrid = static_cast<uint32_t>(std::max<uint64_t>(kXcds, waves / 4));
        pl.fgrid = (pl.nchunks + 255) / 256;
        return pl;
    }
    pl.tpw = static_cast<uint32_t>((ntiles + waves - 1) / waves);
    pl.nchunks = (ntiles + pl.tpw - 1) / pl.tpw;
    pl.tgrid = (pl.nchunks + 3) / 4;
    pl.fgrid = (pl.nchunks + 255) / 256;
    return pl;
}

// Coded kernels exist for sigma <= 8.
template<int KMAX, int S>
constexpr bool coded_kernel() {
    return S <= static_cast<int>(kMaxCodedSigma);
}

template<typename V, int KMAX, bool DOT>
Plan plan_by_sigma(uint32_t sigma, uint32_t ntiles, bool codes) {
    switch (sigma) {
#define DIM_PLAN_CASE(S)                                                                                              \
    case S:                                                                                                           \
        if constexpr (coded_kernel<KMAX, S>())                                                                        \
            if (codes)                                                                                                \
                return make_plan<V, S, KMAX, DOT, true>(ntiles);                                                      \
        return make_plan<V, S, KMAX, DOT, false>(ntiles);
        DIM_FOR_EACH_SIGMA(DIM_PLAN_CASE)
#undef DIM_PLAN_CASE
        default: return Plan{};
    }
}

// nt: single-vector products with the dot stream the matrix non-temporally
// (the other products' choice is fixed by their template arguments).
template<typename V, int S, int KMAX, bool DOT, bool CODES>
void launch_tiles(const TileParams& tp, uint32_t grid, hipStream_t st, bool nt) {
    if constexpr (KMAX == 1 && DOT) {
        if (nt) {
            hipLaunchKernelGGL((k_csr5_tiles<V, S, KMAX, DOT, CODES, true>), dim3(grid), dim3(256), 0, st, tp);
            return;
        }
    }
    hipLaunchKernelGGL((k_csr5_tiles<V, S, KMAX, DOT, CODES>), dim3(grid), dim3(256), 0, st, tp);
}

template<typename V, int KMAX, bool DOT>
void launch_tiles_by_sigma(uint32_t sigma, const TileParams& tp, uint32_t grid, hipStream_t st, bool codes,
                           bool nt) {
    switch (sigma) {
#define DIM_TILES_CASE(S)                                                                                             \
    case S:                                                                                                           \
        if constexpr (coded_kernel<KMAX, S>())                                                                        \
            if (codes) {                                                                                              \
                launch_tiles<V, S, KMAX, DOT, true>(tp, grid, st, nt);                                                \
                break;                                                                                                \
            }                                                                                                         \
        launch_tiles<V, S, KMAX, DOT, false>(tp, grid, st, nt);                                                       \
        break;
        DIM_FOR_EACH_SIGMA(DIM_TILES_CASE)
#undef DIM_TILES_CASE
        default: break;
    }
}

// Whether a solver's product (fused with the dot) should stream the matrix
// non-temporally: its working set, the matrix streams and the solver's four
// vectors (x, r, p, q) of vb-byte values, exceeds the 256 MB Infinity Cache.
bool solve_streams_matrix(const DevFormat& A, size_t vb);

// Launches the tile kernel and the fix-up kernel; records `mid` between them
// when given. Returns the number of dot partials per column and, in *parts,
// where they are: one per fix-up block in fp.dot_blk, or, when a
// single-vector product needs no fix-up, the tile kernel's partials in
// tp.dot_chunk (dot_partials()).
template<typename V, int KMAX>
uint32_t launch_pair(const DevFormat& A, TileParams tp, FixParams fp, bool dot, hipStream_t st, hipEvent_t mid,
                     const double** parts = nullptr) {
    if (parts)
        *parts = fp.dot_blk;
    const uint32_t ntiles = tp.tile_end - tp.tile_begin;
    const bool codes = A.col_code != nullptr;
    // a batch's products over the whole matrix walk its chunk table (chunks
    // that never hold two systems)
    const bool table = A.chunk_tile && tp.tile_begin == 0 && tp.tile_end == A.tiles;
    Plan pl = dot ? plan_by_sigma<V, KMAX, true>(A.sigma, ntiles, codes)
                  : plan_by_sigma<V, KMAX, false>(A.sigma, ntiles, codes);
    if (table) {
        pl.sweep = false;
        pl.tpw = 0;
        pl.nchunks = A.n_chunk_tab;
        pl.tgrid = (pl.nchunks + 3) / 4;
        pl.fgrid = (pl.nchunks + 255) / 256;
        tp.chunk_tile = fp.chunk_tile = A.chunk_tile;
        tp.chunk_sys = A.chunk_sys;
    } else {
        tp.chunk_tile = fp.chunk_tile = nullptr;
        tp.chunk_sys = nullptr;
        tp.sys_done = nullptr;
    }
    tp.tpw = fp.tpw = pl.tpw;
    tp.nchunks = fp.nchunks = pl.nchunks;
    if (!pl.sweep)
        tp.queue = nullptr;
    fp.ndot = table ? pl.nchunks : dot_partials(KMAX, pl.nchunks);
    // Single-vector products finish rows across chunk boundaries by
    // look-ahead and write the empty rows themselves; the fix-up kernel then
    // runs only when some boundary needs it (with a dot, the chunks' partials
    // are then left for the caller to reduce).
    bool fixup = true;
    if (KMAX == 1 && pl.tgrid) {
        tp.lookahead = 1;
        fp.empty_end = fp.empty_begin;
        fixup = table ? A.chunk_tab_fixup
                      : !A.binfo ||
                          A.binfo->needs_fixup(A.binfo_base + tp.tile_begin, A.binfo_base + tp.tile_end, pl.tpw);
    } else {
        tp.lookahead = 0;
        tp.empty_end = tp.empty_begin;
    }
    if (pl.tgrid) {
        if (dot)
            launch_tiles_by_sigma<V, KMAX, true>(A.sigma, tp, pl.tgrid, st, codes,
                                                 KMAX == 1 && solve_streams_matrix(A, sizeof(V)));
        else
            launch_tiles_by_sigma<V, KMAX, false>(A.sigma, tp, pl.tgrid, st, codes, false);
    }
    if (mid)
        (void)hipEventRecord(mid, st);
    if (!fixup) {
        if (!dot)
            return 0; // no fix-up kernel
        if (parts)
            *parts = tp.dot_chunk;
        return fp.ndot;
    }
    const bool per_chunk = table && dot && KMAX == 1;
    fp.dot_rw = per_chunk ? tp.dot_chunk : nullptr;
    if (dot)
        hipLaunchKernelGGL((k_csr5_fixup<V, KMAX, true>), dim3(pl.fgrid), dim3(256), 0, st, fp);
    else
        hipLaunchKernelGGL((k_csr5_fixup<V, KMAX, false>), dim3(pl.fgrid), dim3(256), 0, st, fp);
    if (per_chunk) {
        if (parts)
            *parts = tp.dot_chunk;
        return fp.ndot;
    }
    return pl.fgrid;
}

// Wavefronts of the FP64 single-vector tile kernel (with or without the dot
// epilogue) that the device keeps resident for a format of tile height sigma.
uint64_t resident_tile_waves(uint32_t sigma, bool codes, bool dot);

// Parameters of the single-vector product y = A x with p.q-style dot
// partials over a batch's whole chunk table (look-ahead on, empty rows
// written), for kernels that call csr5_chunk themselves.
TileParams make_product_params(const DevFormat& A, Scratch& S, const void* x, void* y);

// Multi-RHS launchers, one translation unit per k bound (spmv_k*.hip).
uint32_t launch_pair_k2(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, bool dot,
                        hipStream_t st);
uint32_t launch_pair_k4(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, bool dot,
                        hipStream_t st);
uint32_t launch_pair_k8(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, bool dot,
                        hipStream_t st);
uint32_t launch_pair_k16(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, bool dot,
                         hipStream_t st);

} // namespace detail
} // namespace dimgpu
