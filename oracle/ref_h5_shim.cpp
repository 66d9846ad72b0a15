// Thin extern "C" shim over the reference's own HDF5 I/O
// (libs/dim/io/h5.cpp, libs/dim/io/h5/*.cpp, include/dim/io/h5.h), compiled by
// oracle/Makefile straight from the reference tree, against the HDF5 C
// library, into oracle/_ref/libdimref_h5.so.
//
// TEST INFRASTRUCTURE ONLY: it writes matrix files the way the reference's
// store-matrix does and reads them back the way its benchmarks do, so that the
// tests can check files crossing between the two tool chains. It is never part
// of the product.
#include <dim/io/h5.h>
#include <dim/mat/storage_formats/csr.h>
#include <dim/mat/storage_formats/csr5.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <exception>
#include <span>

#include <omp.h>

namespace h5 = dim::io::h5;
using csr5_t = dim::mat::csr5<double>;
using csr_t = dim::mat::csr<double>;

namespace {
csr_t make_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx,
               const double* vals) {
    csr_t::values_t v(vals, vals + nnz);
    csr_t::indices_t rp(row_ptr, row_ptr + rows + 1);
    csr_t::indices_t ci(col_idx, col_idx + nnz);
    return csr_t{dim::mat::dimensions_t{rows, cols}, std::move(v), std::move(rp), std::move(ci)};
}
} // namespace

extern "C" {

// store_matrix -f csr5: csr5::from_csr then h5::store into a new file (or an
// existing one when append) under `group`. 0 on success.
int dimref_h5_store_csr5(const char* path, const char* group, int append, uint32_t rows, uint32_t cols,
                         uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals) {
    try {
        auto file = append ? h5::file_t::open(path, H5F_ACC_RDWR) : h5::file_t::create(path, H5F_ACC_TRUNC);
        auto g = file.create_group(group);
        h5::store(g, csr5_t::from_csr(make_csr(rows, cols, nnz, row_ptr, col_idx, vals)));
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// store_matrix -f csr: write_matlab_compatible.
int dimref_h5_store_matlab(const char* path, const char* group, int append, uint32_t rows, uint32_t cols,
                           uint64_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx, const double* vals) {
    try {
        auto file = append ? h5::file_t::open(path, H5F_ACC_RDWR) : h5::file_t::create(path, H5F_ACC_TRUNC);
        auto g = file.create_group(group);
        h5::write_matlab_compatible(g, make_csr(rows, cols, nnz, row_ptr, col_idx, vals));
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// h5::load_csr5 (h5.cpp:57-74); null on failure.
void* dimref_h5_load_csr5(const char* path, const char* group) {
    try {
        return new csr5_t(h5::load_csr5(path, group));
    } catch (const std::exception&) {
        return nullptr;
    }
}

void dimref_h5_csr5_free(void* h) { delete static_cast<csr5_t*>(h); }

// sizes of a loaded CSR5 matrix: rows, cols, nnz, tiles, tile_desc_offset entries
void dimref_h5_csr5_sizes(void* h, uint64_t* out5) {
    const auto& m = *static_cast<csr5_t*>(h);
    out5[0] = m.dimensions.rows;
    out5[1] = m.dimensions.cols;
    out5[2] = m.vals.size();
    out5[3] = m.csr5_info.tile_count;
    out5[4] = m.csr5_info.tile_desc_offset.size();
}

// the loaded arrays, as stored (vals/col_idx tile-transposed, tile_desc as
// its packed words, 4 per tile)
void dimref_h5_csr5_copy(void* h, uint32_t* row_ptr, uint32_t* tile_ptr, uint32_t* tile_desc_words,
                         uint32_t* tile_desc_offset_ptr, uint32_t* tile_desc_offset, double* vals,
                         uint32_t* col_idx) {
    const auto& m = *static_cast<csr5_t*>(h);
    const auto& info = m.csr5_info;
    std::copy(m.row_ptr.begin(), m.row_ptr.end(), row_ptr);
    for (size_t i = 0; i < info.tile_ptr.size(); ++i)
        tile_ptr[i] = info.tile_ptr[i].raw;
    static_assert(sizeof(csr5_t::tile_descriptor_type) == 16);
    if (!info.tile_desc.empty())
        std::memcpy(tile_desc_words, info.tile_desc.data(), info.tile_desc.size() * 16);
    std::copy(info.tile_desc_offset_ptr.begin(), info.tile_desc_offset_ptr.end(), tile_desc_offset_ptr);
    std::copy(info.tile_desc_offset.begin(), info.tile_desc_offset.end(), tile_desc_offset);
    std::copy(m.vals.begin(), m.vals.end(), vals);
    std::copy(m.col_idx.begin(), m.col_idx.end(), col_idx);
}

// h5::load_csr5_partial (h5.cpp:137-173): part `idx` of `count` of the group's
// tiles, as one rank of the distributed CG loads it; null on failure.
void* dimref_h5_load_csr5_partial(const char* path, const char* group, uint64_t idx, uint64_t count) {
    try {
        const auto file = h5::file_t::open(path, H5F_ACC_RDONLY);
        return new csr5_t(h5::load_csr5_partial(file.open_group(group), h5::partial_identifier_t{idx, count}));
    } catch (const std::exception&) {
        return nullptr;
    }
}

// rows, cols (global), first_row_idx(), last_row_idx() (csr5.h:845-855),
// val_offset, skip_tail, tile_count, loaded elements, loaded row_ptr entries,
// first_tile_uncapped() (csr5.h:845-847)
void dimref_h5_part_info(void* h, uint64_t* out10) {
    const auto& m = *static_cast<csr5_t*>(h);
    out10[0] = m.dimensions.rows;
    out10[1] = m.dimensions.cols;
    out10[2] = m.first_row_idx();
    out10[3] = m.last_row_idx();
    out10[4] = m.val_offset;
    out10[5] = m.skip_tail ? 1 : 0;
    out10[6] = m.csr5_info.tile_count;
    out10[7] = m.vals.size();
    out10[8] = m.row_ptr.size();
    out10[9] = m.first_tile_uncapped() ? 1 : 0;
}

// the loaded slabs: row_ptr (absolute element indices), tile_ptr (raw words),
// vals/col_idx (as stored, tile-transposed)
void dimref_h5_part_copy(void* h, uint32_t* row_ptr, uint32_t* tile_ptr, double* vals, uint32_t* col_idx) {
    const auto& m = *static_cast<csr5_t*>(h);
    std::copy(m.row_ptr.begin(), m.row_ptr.end(), row_ptr);
    for (size_t i = 0; i < m.csr5_info.tile_ptr.size(); ++i)
        tile_ptr[i] = m.csr5_info.tile_ptr[i].raw;
    std::copy(m.vals.begin(), m.vals.end(), vals);
    std::copy(m.col_idx.begin(), m.col_idx.end(), col_idx);
}

// csr5::spmv<spmv_strategy::partial> (csr5.h:818-829) of a loaded part with
// `nthreads` OpenMP threads: y_block (last_row_idx() - first_row_idx() + 1
// entries) is zero-filled and the calibrator is fresh, as the reference
// requires (its tile loop adds into y). x has the global column count.
int dimref_h5_part_spmv(void* h, int nthreads, const double* x, double* y_block) {
    try {
        const auto& m = *static_cast<csr5_t*>(h);
        omp_set_num_threads(nthreads);
        const size_t n = m.last_row_idx() - m.first_row_idx() + 1;
        std::fill(y_block, y_block + n, 0.0);
        auto calibrator = m.create_calibrator();
        m.spmv<csr5_t::spmv_strategy::partial>(
          csr5_t::spmv_data_t{std::span<const double>(x, m.dimensions.cols), std::span<double>(y_block, n), calibrator});
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// read_matlab_compatible (h5.h:87-97); null on failure.
void* dimref_h5_read_matlab(const char* path, const char* group) {
    try {
        return new csr_t(h5::read_matlab_compatible<dim::mat::cache_aligned_vector>(path, group));
    } catch (const std::exception&) {
        return nullptr;
    }
}

void dimref_h5_csr_free(void* h) { delete static_cast<csr_t*>(h); }

void dimref_h5_csr_sizes(void* h, uint64_t* out3) {
    const auto& m = *static_cast<csr_t*>(h);
    out3[0] = m.dimensions.rows;
    out3[1] = m.dimensions.cols;
    out3[2] = m.values.size();
}

void dimref_h5_csr_copy(void* h, uint32_t* row_ptr, uint32_t* col_idx, double* vals) {
    const auto& m = *static_cast<csr_t*>(h);
    std::copy(m.row_start_offsets.begin(), m.row_start_offsets.end(), row_ptr);
    std::copy(m.col_indices.begin(), m.col_indices.end(), col_idx);
    std::copy(m.values.begin(), m.values.end(), vals);
}

} // extern "C"
