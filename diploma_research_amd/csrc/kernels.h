// Device-side view of the GPU CSR5 format and the kernel launch entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

namespace dimgpu {

// Host copy of the tile flags, used to decide whether a product needs the
// fix-up kernel: with chunks of `tpw` tiles, a row that crosses into a chunk
// is finished by the earlier chunk (look-ahead) unless it began before that
// chunk's last tile or runs past the later chunk's first tile.
class BoundaryInfo {
  public:
    void assign(std::vector<uint8_t> flags) {
        std::lock_guard<std::mutex> g(m_);
        flags_ = std::move(flags);
        memo_valid_ = false;
    }
    // tiles [tb, te) of the flag array, chunks of tpw tiles
    bool needs_fixup(uint32_t tb, uint32_t te, uint32_t tpw) const;
    // the same for chunks at the boundaries chunk_tile[0..n]
    bool needs_fixup_table(const std::vector<uint32_t>& chunk_tile) const;

  private:
    std::vector<uint8_t> flags_;
    mutable std::mutex m_;
    mutable bool memo_valid_ = false;
    mutable uint32_t memo_tb_ = 0, memo_te_ = 0, memo_tpw_ = 0;
    mutable bool memo_need_ = true;
};

// Device pointers of one matrix (see HostFormat for the layout).
struct DevFormat {
    uint32_t rows = 0, cols = 0, sigma = 16, tiles = 0;
    uint64_t nnz = 0, padded = 0;
    uint32_t* tile_ptr = nullptr;
    void* bit_flag = nullptr; // u16 words for sigma <= 16, u32 otherwise
    uint8_t* tile_flags = nullptr;
    uint32_t* tile_desc_offset_ptr = nullptr;
    uint32_t* tile_desc_offset = nullptr;
    uint32_t* col_idx = nullptr;
    double* vals64 = nullptr;
    float* vals32 = nullptr;
    uint32_t* empty_rows = nullptr;
    uint32_t n_empty = 0;
    bool any_dirty = false;
    // column codes (HostFormat::col_code / col_dict), or null
    uint32_t* col_code = nullptr;
    int32_t* col_dict = nullptr; // kMaxDiagonals entries
    // host-side tile flags of the whole matrix; tile t of this view is tile
    // binfo_base + t there
    const BoundaryInfo* binfo = nullptr;
    uint32_t binfo_base = 0;
    // Block-diagonal batches: the chunk table (chunk c covers tiles
    // [chunk_tile[c], chunk_tile[c + 1]) of system chunk_sys[c]; chunks never
    // hold two systems), used by products over the whole matrix, and whether
    // some chunk boundary needs the fix-up kernel. With column codes col_dict
    // holds kMaxDiagonals entries per system.
    const uint32_t* chunk_tile = nullptr;
    const uint32_t* chunk_sys = nullptr;
    uint32_t n_chunk_tab = 0;
    bool chunk_tab_fixup = true;
};

// The multi-RHS products' ordered sweep (TileParams::queue): one counter
// pair per XCD, 128 bytes apart, zero between launches (the last wavefront
// of each range clears its pair).
constexpr uint32_t kXcds = 8;
constexpr uint32_t kQueueStride = 32;
constexpr uint32_t kQueueWords = kXcds * kQueueStride;

// Per-matrix scratch: tile carries and reduction partials (the GPU
// counterpart of the reference's calibrator, include/dim/mat/storage_formats/
// csr5/calibrator.h:14-31). Sized for up to kMaxK right-hand sides.
constexpr int kMaxK = 16;
struct Scratch {
    void* head = nullptr;          // chunks * kMaxK * 8 bytes (chunks <= tiles)
    void* tail = nullptr;          // chunks * kMaxK * 8 bytes
    uint8_t* chunk_flags = nullptr; // chunks
    double* dot_chunk = nullptr;   // chunks * kMaxK
    double* dot_blk = nullptr;     // ceil(chunks/256) * kMaxK
    uint32_t* queue = nullptr;     // kQueueWords, zero between products
};

// y = A x over tiles [tile_begin, tile_end). x and y are n x k, row-major.
// row_base: y row index offset (0 for the whole matrix; first row of a block).
// When `dot` is set (square matrices, whole matrix only) the per-column
// products x[row]*y[row] are left as partials, column j of partial i at
// parts[i*k + j]: *dot_parts (if given) receives parts (scratch.dot_blk, one
// partial per fix-up block, or scratch.dot_chunk, one per chunk, when a
// single-vector product needs no fix-up kernel) and *dot_blocks (if given)
// the number of partials.
// sys_done (single-vector products of a batch over the whole matrix; may be
// null): the systems' stop flags; the chunks of stopped systems leave their
// rows of y as they are.
// `done` (may be null): a device flag; when *done != 0 the kernels return at
// once (used by solver launch sequences).
// ldx: 0 for row-major n x k blocks (element (i, j) at i*k + j); otherwise the
// blocks are column-major with element (i, j) of x at i + j*ldx and of y at
// i + j*ldy (ldy = 0: same as ldx).
hipError_t launch_spmv(const DevFormat& A, Scratch& S, bool f32, int k, const void* x, void* y,
                       uint32_t tile_begin, uint32_t tile_end, uint32_t row_base, uint32_t empty_begin,
                       uint32_t empty_end, bool dot, const int* done, hipStream_t stream,
                       uint32_t* dot_blocks = nullptr, uint64_t ldx = 0, uint64_t ldy = 0,
                       const double** dot_parts = nullptr, const int* sys_done = nullptr);

// Same as launch_spmv over the whole matrix with events recorded before the
// tile kernel, between the two kernels and after the fix-up kernel.
hipError_t launch_spmv_events(const DevFormat& A, Scratch& S, bool f32, const void* x, void* y, hipEvent_t e0,
                              hipEvent_t e1, hipEvent_t e2, hipStream_t stream, bool* fixup_launched = nullptr);

// y = A^T x (x has rows entries, y has cols entries); y is zeroed first.
hipError_t launch_spmv_transpose(const DevFormat& A, bool f32, const void* x, void* y, hipStream_t stream,
                                 const int* done = nullptr);

// y[i] = sum over the column slices s = 0..slices-1, in that order, of the
// partial product of (s, i): the partials of the pairs whose bit s*rows + i
// is set in `present`, stored at the rank of that bit (rank[w] = set bits
// before word w); the other pairs count as zero.
hipError_t launch_slice_sum(bool f32, const void* ystack, const uint32_t* present, const uint32_t* rank,
                            uint32_t rows, uint32_t slices, void* y, hipStream_t stream);

// n x k blocks from row-major to column-major (to_cols) or back.
hipError_t launch_block_transpose(bool f32, bool to_cols, size_t n, int k, const void* src, void* dst,
                                  hipStream_t stream);

// Scatters CSR-order values into the tile layout (FP64, and FP32 if kept).
hipError_t launch_set_values(const DevFormat& A, const double* vals_csr, hipStream_t stream);

// out[e] = scale * u[row(e)] * v[col(e)] for every stored element e, in CSR order.
hipError_t launch_sampled_outer(const DevFormat& A, double scale, const double* u, const double* v, double* out,
                                hipStream_t stream);

} // namespace dimgpu
