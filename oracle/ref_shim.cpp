// Thin extern "C" shim over the reference's own header-only CSR5 code
// (include/dim/mat/storage_formats/csr5.h), compiled by oracle/Makefile
// straight from the reference tree into oracle/_ref/libdimref.so.
//
// TEST INFRASTRUCTURE ONLY: it pins the oracle to the reference and serves as
// the reference CPU baseline. It is never part of the product.
#include <dim/mat/storage_formats/coo.h>
#include <dim/mat/storage_formats/csr.h>
#include <dim/mat/storage_formats/csr5.h>

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

using csr5_t = dim::mat::csr5<double>;
using csr_t = dim::mat::csr<double>;
using coo_t = dim::mat::coo<double>;

extern "C" {

// csr5::from_csr (csr5.h:791-803)
void* dimref_csr5_from_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                           const uint32_t* col_idx, const double* vals) {
    csr_t::values_t v(vals, vals + nnz);
    csr_t::indices_t rp(row_ptr, row_ptr + rows + 1);
    csr_t::indices_t ci(col_idx, col_idx + nnz);
    csr_t csr{dim::mat::dimensions_t{rows, cols}, std::move(v), std::move(rp), std::move(ci)};
    return new csr5_t(csr5_t::from_csr(std::move(csr)));
}

void dimref_csr5_free(void* h) { delete static_cast<csr5_t*>(h); }

uint64_t dimref_csr5_tile_count(void* h) { return static_cast<csr5_t*>(h)->csr5_info.tile_count; }
uint64_t dimref_csr5_tile_desc_offset_size(void* h) {
    return static_cast<csr5_t*>(h)->csr5_info.tile_desc_offset.size();
}
uint32_t dimref_csr5_tail_partition_start(void* h) { return static_cast<csr5_t*>(h)->tail_partition_start(); }

// Copies the converted arrays out; tile_desc is copied as its raw packed words
// (4 per tile), exactly as h5::store writes them (h5.cpp:45-55).
void dimref_csr5_copy(void* h, uint32_t* tile_ptr, uint32_t* tile_desc_words, uint32_t* tile_desc_offset_ptr,
                      uint32_t* tile_desc_offset, double* vals, uint32_t* col_idx) {
    const auto& m = *static_cast<csr5_t*>(h);
    const auto& info = m.csr5_info;
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

// csr5::spmv (csr5.h:818-829) called the way it is correct: y zero-filled and a
// fresh calibrator, with `nthreads` OpenMP threads. Returns 1 when the
// reference throws std::invalid_argument.
int dimref_csr5_spmv(void* h, int nthreads, const double* x, uint64_t x_len, double* y, uint64_t y_len) {
    const auto& m = *static_cast<csr5_t*>(h);
    omp_set_num_threads(nthreads);
    std::fill(y, y + y_len, 0.0);
    auto calibrator = m.create_calibrator();
    try {
        m.spmv({.x = std::span<const double>(x, x_len), .y = std::span<double>(y, y_len), .calibrator = calibrator});
    } catch (const std::invalid_argument&) {
        return 1;
    }
    return 0;
}

// The reference's benchmark protocol (benchmarks/csr5/dim/main.cpp:58-74):
// `warmup` products, then `runs` timed products, each timed alone with
// omp_get_wtime. Before every product y is zero-filled and the calibrator is
// fresh (outside the timed region), which the reference's own loop omits
// (its stale-state defect). times[runs] receives the seconds of each product.
void dimref_csr5_time_spmv(void* h, int nthreads, const double* x, uint64_t x_len, double* y, uint64_t y_len,
                           int warmup, int runs, double* times) {
    const auto& m = *static_cast<csr5_t*>(h);
    omp_set_num_threads(nthreads);
    for (int i = 0; i < warmup + runs; ++i) {
        std::fill(y, y + y_len, 0.0);
        auto calibrator = m.create_calibrator();
        const double t0 = omp_get_wtime();
        m.spmv({.x = std::span<const double>(x, x_len), .y = std::span<double>(y, y_len), .calibrator = calibrator});
        const double t1 = omp_get_wtime();
        if (i >= warmup)
            times[i - warmup] = t1 - t0;
    }
}

// Textbook CG (thesis.tex:2133-2160) in the step order of
// benchmarks/conjugate_gradient/dim_sp/main.cpp:95-127 (As = A s; alpha =
// r.r / s.As; x += alpha s; r -= alpha As; r.r; stop test; beta; s = r + beta
// s) with the reference's csr5::spmv for every product (y zero-filled, fresh
// calibrator) and OpenMP loops for the vector steps (the reference's vec.h
// needs a parallel STL backend this image does not have). Stops when
// sqrt(r.r) / ||b|| <= rtol or after maxit iterations; x holds x0.
// out[0..3]: iterations, converged, recursive relres, true relres (FP64
// ||b - A x|| / ||b|| recomputed at exit); out[4]: seconds of the iterations.
void dimref_cg(void* h, int nthreads, const double* b, double* x, double rtol, uint32_t maxit, double* out) {
    const auto& m = *static_cast<csr5_t*>(h);
    omp_set_num_threads(nthreads);
    const int64_t n = m.dimensions.rows;
    std::vector<double> r(n), s(n), q(n);
    auto product = [&](const double* in, double* y) {
        std::fill(y, y + n, 0.0);
        auto calibrator = m.create_calibrator();
        m.spmv({.x = std::span<const double>(in, m.dimensions.cols), .y = std::span<double>(y, n),
                .calibrator = calibrator});
    };
    auto dot = [&](const double* a, const double* c) {
        double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (int64_t i = 0; i < n; ++i)
            acc += a[i] * c[i];
        return acc;
    };
    const double t0 = omp_get_wtime();
    product(x, q.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        r[i] = b[i] - q[i];
        s[i] = r[i];
    }
    const double bnorm = std::sqrt(dot(b, b));
    const double scale = bnorm > 0.0 ? bnorm : 1.0;
    double rr = dot(r.data(), r.data());
    uint32_t it = 0;
    bool conv = std::sqrt(rr) / scale <= rtol;
    while (!conv && it < maxit) {
        product(s.data(), q.data());
        const double alpha = rr / dot(s.data(), q.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * s[i];
            r[i] -= alpha * q[i];
        }
        const double rr_new = dot(r.data(), r.data());
        ++it;
        if (std::sqrt(rr_new) / scale <= rtol) {
            conv = true;
            rr = rr_new;
            break;
        }
        const double beta = rr_new / rr;
        rr = rr_new;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            s[i] = r[i] + beta * s[i];
    }
    const double t1 = omp_get_wtime();
    product(x, q.data());
    double tr = 0.0;
#pragma omp parallel for reduction(+ : tr) schedule(static)
    for (int64_t i = 0; i < n; ++i)
        tr += (b[i] - q[i]) * (b[i] - q[i]);
    out[0] = it;
    out[1] = conv ? 1.0 : 0.0;
    out[2] = std::sqrt(rr) / scale;
    out[3] = std::sqrt(tr) / scale;
    out[4] = t1 - t0;
}

// csr5::spmv exactly as the reference's own runners call it: y and the
// calibrator are whatever the caller kept from the previous call
// (benchmarks/csr5/dim/main.cpp:197-209). Used to document the stale-state
// defect and to time the reference the way its benchmark does.
struct dimref_session {
    csr5_t* m;
    csr5_t::calibrator_type calibrator;
};
void* dimref_session_new(void* h, int nthreads) {
    omp_set_num_threads(nthreads);
    auto* m = static_cast<csr5_t*>(h);
    return new dimref_session{m, m->create_calibrator()};
}
void dimref_session_free(void* s) { delete static_cast<dimref_session*>(s); }
int dimref_session_spmv(void* s, const double* x, uint64_t x_len, double* y, uint64_t y_len) {
    auto* ss = static_cast<dimref_session*>(s);
    try {
        ss->m->spmv(
          {.x = std::span<const double>(x, x_len), .y = std::span<double>(y, y_len), .calibrator = ss->calibrator});
    } catch (const std::invalid_argument&) {
        return 1;
    }
    return 0;
}

// csr::from_coo (csr.h:42-77): out arrays sized rows+1 / nnz_out (nnz_out is
// nnz, or up to 2*nnz for a symmetric COO). Returns the number of stored
// entries.
uint64_t dimref_csr_from_coo(uint32_t rows, uint32_t cols, uint64_t nnz, int symmetric, const uint32_t* ri,
                             const uint32_t* ci, const double* v, uint32_t* row_ptr, uint32_t* col_idx,
                             double* vals) {
    coo_t coo{dim::mat::dimensions_t{rows, cols}, nnz, symmetric != 0};
    std::copy(ri, ri + nnz, coo.row_indices.begin());
    std::copy(ci, ci + nnz, coo.col_indices.begin());
    std::copy(v, v + nnz, coo.values.begin());
    auto csr = csr_t::from_coo(coo);
    std::copy(csr.row_start_offsets.begin(), csr.row_start_offsets.end(), row_ptr);
    std::copy(csr.col_indices.begin(), csr.col_indices.end(), col_idx);
    std::copy(csr.values.begin(), csr.values.end(), vals);
    return csr.values.size();
}

} // extern "C"
