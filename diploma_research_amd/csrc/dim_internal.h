// Internal declarations shared by the host format builder, the HIP kernels
// and the C ABI of libdim_gpu.so.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dimgpu {

// One CSR5 tile column is one lane of a 64-wide wavefront.
constexpr uint32_t kOmega = 64;
// tile_ptr MSB: the tile's rows include an empty row, so its segments map to
// rows through tile_desc_offset (same convention as the reference,
// include/dim/mat/storage_formats/csr5.h:156-169, 342-397).
constexpr uint32_t kDirty = 0x80000000u;
// tile_flags bits.
constexpr uint8_t kStartsRow = 1; // the tile's first element is the first element of a row
constexpr uint8_t kMultiSeg = 2;  // the tile holds more than one row segment

// Tile heights with compiled kernels. Small and odd values let the autotuner
// match sigma to short, regular rows (a 7-point stencil at sigma = 7 keeps a
// lane on one row, so the x gathers of a wavefront step are contiguous).
#define DIM_FOR_EACH_SIGMA(X) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28) X(32)
inline bool sigma_supported(uint32_t s) {
#define DIM_SIGMA_CASE(S) if (s == S) return true;
    DIM_FOR_EACH_SIGMA(DIM_SIGMA_CASE)
#undef DIM_SIGMA_CASE
    return false;
}

// Host image of the GPU CSR5 format (omega = 64, sigma from DIM_FOR_EACH_SIGMA).
//
//  * tiles = ceil(nnz / (64*sigma)); the last tile is padded (val 0, col 0)
//    instead of keeping the reference's separate CSR tail partition.
//  * vals/col_idx: inside tile t the CSR element t*64*sigma + c*sigma + r sits
//    at t*64*sigma + r*64 + c, so step r of a wavefront is one contiguous
//    512-byte (f64) load.
//  * bit_flag[t*64 + c]: bit r set iff that element starts a row.
//  * tile_ptr[t]: row holding the tile's first element (MSB = dirty);
//    tile_ptr[tiles] = rows.
//  * tile_desc_offset: for dirty tiles, the row of each segment relative to
//    tile_ptr (segment 0 begins at the tile's first element).
//  * empty_rows: rows without non-zeros; the kernels write 0 there.
struct HostFormat {
    uint32_t rows = 0, cols = 0, sigma = 16, tiles = 0;
    uint64_t nnz = 0, padded = 0;
    std::vector<uint32_t> tile_ptr;
    std::vector<uint32_t> bit_flag; // widened; narrowed to u16 on upload for sigma <= 16
    std::vector<uint8_t> tile_flags;
    std::vector<uint32_t> tile_desc_offset_ptr;
    std::vector<uint32_t> tile_desc_offset;
    std::vector<uint32_t> col_idx;
    std::vector<double> vals;
    std::vector<uint32_t> empty_rows;
    bool any_dirty = false;
    // Column codes (see build_col_codes): empty unless every element lies on
    // one of few diagonals and no row is empty.
    std::vector<int32_t> col_dict;  // the diagonals col - row, ascending
    std::vector<uint32_t> col_code; // T*64 words, see build_col_codes
};

// Column codes: for matrices whose elements lie on few diagonals (col - row
// takes few values, as in stencil operators) and have no empty rows, each
// lane's column of a tile is also stored as one 32-bit word: bits [0, 24)
// hold the sigma codes of its elements, code_bits(sigma) bits each, an index
// into the table of diagonals; bits [24, 24 + sigma) hold its row-start flags
// (the bit_flag word). The product then rebuilds col = row + table[code] from
// the row starts instead of streaming 4-byte column indices and the flags.
constexpr uint32_t kMaxDiagonals = 16;
constexpr uint32_t kMaxCodedSigma = 8;
constexpr uint32_t kCodeFlagShift = 24;
constexpr uint32_t code_bits(uint32_t sigma) { return sigma <= 6 ? 4 : 3; }
void build_col_codes(HostFormat& f, const uint32_t* row_ptr, const uint32_t* col_idx);

// Builds the format from CSR. Returns an empty string or an error message.
std::string build_format(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const double* vals, uint32_t sigma, HostFormat& out);

// Tile heights the autotuner times for a matrix with `rows` rows and `nnz`
// stored elements.
std::vector<uint32_t> sigma_candidates(uint32_t rows, uint64_t nnz);

// Validates CSR structure (monotone row_ptr, row_ptr[0]==0, row_ptr[rows]==nnz,
// column indices < cols). Returns an empty string or an error message.
std::string validate_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx);

} // namespace dimgpu
