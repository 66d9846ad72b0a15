// CSR -> GPU CSR5 (omega = 64) conversion on the host.
//
// Follows the reference conversion (include/dim/mat/storage_formats/csr5.h:
// generate_tile_pointer 411-431, set_bit_flag 323-340, generate_tile_offset
// 363-397, transpose 463-493) with two changes of layout for the GPU:
// omega is the wavefront width, and the last partial tile is padded instead
// of being left to a scalar tail pass. The per-column y_offset and
// scansum_offset of the reference are not stored: the kernel derives both in
// registers from the bit flags with a wavefront prefix sum and a ballot.
#include "dim_internal.h"

#include <algorithm>
#include <climits>
#include <mutex>
#include <thread>

namespace dimgpu {

namespace {

// index of the last element <= v in a[0..n) (csr5.h:194-198)
inline uint64_t upper_bound_idx(const uint32_t* a, uint64_t n, uint64_t v) {
    return static_cast<uint64_t>(std::upper_bound(a, a + n, v, [](uint64_t x, uint32_t y) { return x < y; }) - a) - 1;
}

template<typename F>
void parallel_for(uint64_t n, F&& f) {
    unsigned hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 4096 || hw == 1) {
        f(uint64_t{0}, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + hw - 1) / hw;
    for (unsigned i = 0; i < hw; ++i) {
        uint64_t b = i * chunk, e = std::min(n, b + chunk);
        if (b >= e)
            break;
        th.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : th)
        t.join();
}

} // namespace

std::string validate_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx) {
    if (!row_ptr)
        return "row_ptr is null";
    if (nnz && !col_idx)
        return "col_idx is null";
    if (rows >= kDirty)
        return "row count must be below 2^31 (the tile pointer's top bit is the dirty flag)";
    if (nnz >= (uint64_t{1} << 32))
        return "non-zero count must fit in 32 bits";
    if (row_ptr[0] != 0)
        return "row_ptr[0] must be 0";
    if (row_ptr[rows] != nnz)
        return "row_ptr[rows] must equal nnz";
    for (uint64_t r = 0; r < rows; ++r)
        if (row_ptr[r + 1] < row_ptr[r])
            return "row_ptr must be non-decreasing";
    for (uint64_t k = 0; k < nnz; ++k)
        if (col_idx[k] >= cols)
            return "column index out of range";
    return {};
}

std::string build_format(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const double* vals, uint32_t sigma, HostFormat& f) {
    if (!sigma_supported(sigma))
        return "sigma must be one of 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32";
    f = HostFormat{};
    f.rows = rows;
    f.cols = cols;
    f.sigma = sigma;
    f.nnz = nnz;
    const uint64_t tile = uint64_t{kOmega} * sigma;
    const uint64_t T = (nnz + tile - 1) / tile;
    if (T >= kDirty)
        return "too many tiles";
    f.tiles = static_cast<uint32_t>(T);
    f.padded = T * tile;
    const uint64_t nrp = uint64_t{rows} + 1;

    // tile_ptr (generate_tile_pointer, csr5.h:411-431)
    f.tile_ptr.assign(T + 1, 0);
    parallel_for(T + 1, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t)
            f.tile_ptr[t] = static_cast<uint32_t>(upper_bound_idx(row_ptr, nrp, std::min(t * tile, nnz)));
    });
    f.tile_ptr[T] = rows; // the padded last tile ends the matrix
    // dirty marks: an empty row among the tile's rows [start, stop]
    std::vector<uint8_t> dirty(T, 0);
    parallel_for(T, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            uint32_t start = f.tile_ptr[t], stop = std::min<uint32_t>(f.tile_ptr[t + 1], rows);
            for (uint32_t r = start; r < stop; ++r)
                if (row_ptr[r] == row_ptr[r + 1]) {
                    dirty[t] = 1;
                    break;
                }
        }
    });

    // bit flags: each non-empty row sets the bit of its first element
    f.bit_flag.assign(T * kOmega, 0);
    parallel_for(rows, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e; ++r) {
            uint64_t idx = row_ptr[r];
            if (row_ptr[r + 1] == idx)
                continue;
            uint64_t t = idx / tile, c = (idx / sigma) % kOmega, bit = idx % sigma;
            // rows in one thread's range never share a tile column word with
            // another thread's rows except at range edges; use an atomic OR.
            __atomic_fetch_or(&f.bit_flag[t * kOmega + c], 1u << bit, __ATOMIC_RELAXED);
        }
    });

    // tile flags and segment counts
    f.tile_flags.assign(T + 1, 0);
    std::vector<uint32_t> nseg(T, 0);
    parallel_for(T, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            uint32_t n = 0;
            for (uint32_t c = 0; c < kOmega; ++c)
                n += static_cast<uint32_t>(__builtin_popcount(f.bit_flag[t * kOmega + c] | (c == 0 ? 1u : 0u)));
            nseg[t] = n;
            uint8_t fl = 0;
            if (f.bit_flag[t * kOmega] & 1u)
                fl |= kStartsRow;
            if (n > 1)
                fl |= kMultiSeg;
            f.tile_flags[t] = fl;
        }
    });
    f.tile_flags[T] = kStartsRow; // sentinel: nothing continues past the last tile

    // segment -> row maps of dirty tiles (generate_tile_offset, csr5.h:363-397)
    f.tile_desc_offset_ptr.assign(T + 1, 0);
    uint64_t total = 0;
    for (uint64_t t = 0; t < T; ++t) {
        f.tile_desc_offset_ptr[t] = static_cast<uint32_t>(total);
        if (dirty[t])
            total += nseg[t];
    }
    f.tile_desc_offset_ptr[T] = static_cast<uint32_t>(total);
    f.any_dirty = total != 0;
    f.tile_desc_offset.assign(std::max<uint64_t>(total, 1), 0);
    parallel_for(T, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            if (!dirty[t])
                continue;
            f.tile_ptr[t] |= kDirty;
            const uint32_t start = f.tile_ptr[t] & ~kDirty;
            uint32_t k = f.tile_desc_offset_ptr[t];
            for (uint32_t c = 0; c < kOmega; ++c) {
                uint32_t bits = f.bit_flag[t * kOmega + c] | (c == 0 ? 1u : 0u);
                for (uint32_t r = 0; r < sigma; ++r) {
                    if (!((bits >> r) & 1u))
                        continue;
                    uint64_t g = t * tile + uint64_t{c} * sigma + r;
                    f.tile_desc_offset[k++] = static_cast<uint32_t>(upper_bound_idx(row_ptr, nrp, g) - start);
                }
            }
        }
    });

    // transposed, padded value and column arrays (transpose, csr5.h:463-493)
    f.vals.assign(f.padded, 0.0);
    f.col_idx.assign(f.padded, 0);
    parallel_for(T, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            const uint64_t base = t * tile;
            for (uint32_t c = 0; c < kOmega; ++c)
                for (uint32_t r = 0; r < sigma; ++r) {
                    uint64_t src = base + uint64_t{c} * sigma + r;
                    if (src >= nnz)
                        continue;
                    uint64_t dst = base + uint64_t{r} * kOmega + c;
                    f.vals[dst] = vals ? vals[src] : 0.0;
                    f.col_idx[dst] = col_idx[src];
                }
        }
    });

    for (uint64_t r = 0; r < rows; ++r)
        if (row_ptr[r] == row_ptr[r + 1])
            f.empty_rows.push_back(static_cast<uint32_t>(r));
    build_col_codes(f, row_ptr, col_idx);
    return {};
}

void build_col_codes(HostFormat& f, const uint32_t* row_ptr, const uint32_t* col_idx) {
    f.col_dict.clear();
    f.col_code.clear();
    // (products of coded matrices address x with 32-bit byte offsets)
    if (f.sigma > kMaxCodedSigma || f.any_dirty || !f.nnz || !f.empty_rows.empty() ||
        uint64_t{f.cols} * sizeof(double) > (uint64_t{1} << 32))
        return;
    // the diagonals, gathered per thread range and merged; more than
    // kMaxDiagonals (or one beyond 32-bit range) means no codes
    std::vector<std::vector<int64_t>> parts;
    std::mutex m;
    bool too_many = false;
    parallel_for(f.rows, [&](uint64_t b, uint64_t e) {
        std::vector<int64_t> d;
        for (uint64_t r = b; r < e && d.size() <= kMaxDiagonals; ++r)
            for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                const int64_t v = int64_t{col_idx[k]} - static_cast<int64_t>(r);
                if (std::find(d.begin(), d.end(), v) == d.end()) {
                    d.push_back(v);
                    if (d.size() > kMaxDiagonals)
                        break;
                }
            }
        std::lock_guard<std::mutex> g(m);
        if (d.size() > kMaxDiagonals)
            too_many = true;
        parts.push_back(std::move(d));
    });
    if (too_many)
        return;
    std::vector<int64_t> all;
    for (auto& d : parts)
        all.insert(all.end(), d.begin(), d.end());
    std::sort(all.begin(), all.end());
    all.erase(std::unique(all.begin(), all.end()), all.end());
    const uint32_t cb = code_bits(f.sigma);
    if (all.size() > (1u << cb) || all.front() < INT32_MIN || all.back() > INT32_MAX)
        return;
    f.col_dict.assign(all.begin(), all.end());
    // code of every element (padding: code 0), packed per tile column with
    // the column's row-start flags
    const uint64_t tile = uint64_t{kOmega} * f.sigma;
    std::vector<uint8_t> code(f.nnz);
    parallel_for(f.rows, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e; ++r)
            for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                const int64_t v = int64_t{col_idx[k]} - static_cast<int64_t>(r);
                code[k] = static_cast<uint8_t>(std::lower_bound(all.begin(), all.end(), v) - all.begin());
            }
    });
    f.col_code.assign(uint64_t{f.tiles} * kOmega, 0);
    parallel_for(uint64_t{f.tiles} * kOmega, [&](uint64_t b, uint64_t e) {
        for (uint64_t w = b; w < e; ++w) {
            const uint64_t first = (w / kOmega) * tile + (w % kOmega) * f.sigma;
            uint32_t word = f.bit_flag[w] << kCodeFlagShift;
            for (uint32_t r = 0; r < f.sigma && first + r < f.nnz; ++r)
                word |= uint32_t{code[first + r]} << (cb * r);
            f.col_code[w] = word;
        }
    });
}

} // namespace dimgpu
