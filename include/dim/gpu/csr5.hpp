// dim/gpu/csr5.hpp - C++ surface of the MI355X CSR5 path over the C ABI
// (dim_gpu.h), shaped after the reference's dim::mat::csr5
// (include/dim/mat/storage_formats/csr5.h:254-880) so that host code written
// against the reference keeps its names:
//
//   reference                                   here
//   csr5<double>::from_csr(csr)                 csr5::from_csr(csr) (host CSR in, device format out)
//   csr5::create_calibrator()                   csr5::create_calibrator() (a token: carries live on the device)
//   csr5::spmv<absolute>({x, y, calibrator})    csr5::spmv({x, y, calibrator}) with DEVICE spans, or
//                                               csr5::spmv_host({x, y, calibrator}) with host spans
//   csr5::spmv<partial>(...) on a part from     the same, on a part from dim::io::h5::load_csr5_partial
//     h5::load_csr5_partial                       (dim/gpu/h5.hpp), y holding last_row_idx() -
//                                                 first_row_idx() + 1 entries; or csr5::spmv_block(
//                                                 tile_begin, tile_end, x, y_block) on a whole matrix
//   csr5::tail_partition_start(), first/last_row_idx(), non_zero_count(), dimensions
//   conjugate_gradient(A, b, x)                 dim::gpu::conjugate_gradient(A, b, x, opts)
//
// dim::mat::csr5<V, Sigma, Omega, U, Storage>, dim::mat::csr<V> and
// dim::mat::dimensions_t name the same types, so that code spelled with the
// reference's namespace compiles unchanged (the GPU tile is omega = 64 with
// sigma from the autotuner, so the Sigma/Omega/U/Storage arguments are
// accepted and not used; V must be double).
//
// Errors are thrown as in the reference: std::invalid_argument for bad
// arguments (csr5.h:820-821), std::runtime_error for everything else. Header
// only; link libdim_gpu.so.
#ifndef DIM_GPU_CSR5_HPP
#define DIM_GPU_CSR5_HPP

#include "../../dim_gpu.h"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace dim::gpu {

inline void check(dim_status st) {
    if (st == DIM_SUCCESS)
        return;
    const std::string msg = std::string(dim_status_string(st)) + ": " + dim_last_error();
    if (st == DIM_INVALID_ARGUMENT)
        throw std::invalid_argument{msg};
    throw std::runtime_error{msg};
}

// dim::mat::dimensions_t (include/dim/mat/storage_formats/base.h)
struct dimensions_t {
    uint32_t rows;
    uint32_t cols;
};

// dim::mat::csr<double> (include/dim/mat/storage_formats/csr.h:19-78), host side
struct csr {
    dimensions_t dimensions{};
    std::vector<double> values;
    std::vector<uint32_t> row_start_offsets;
    std::vector<uint32_t> col_indices;
};

// The reference keeps one carry per OpenMP thread in a calibrator_t
// (csr5/calibrator.h:14-31) that the caller passes to every spmv. The GPU
// path keeps its carries in device scratch owned by the matrix, so the
// calibrator is an empty token kept for source compatibility; as with the
// reference's calibrator, products on one matrix must not overlap.
struct calibrator_t {};

// spmv_data_t (csr5.h:280-285) over DEVICE memory
template<typename V>
struct spmv_data_t {
    std::span<const V> x;
    std::span<V> y;
    calibrator_t& calibrator;
};

enum class spmv_strategy { absolute, partial };

struct csr5_options {
    uint32_t sigma = 0;     // 0: the built-in autotuner
    bool keep_f32 = false;  // keep FP32 values for FP32 products / mixed-precision CG
};

class csr5 {
  public:
    using value_type = double;
    using spmv_strategy = ::dim::gpu::spmv_strategy;
    using calibrator_type = calibrator_t;
    using spmv_data_t = ::dim::gpu::spmv_data_t<double>;

    // One rank's share loaded by h5::load_csr5_partial: the rows of the
    // whole matrix that its products write, and where its elements start.
    struct part_t {
        uint32_t first_row = 0, n_rows = 0;
        uint64_t val_offset = 0;
        bool skip_tail = false;
        bool first_tile_uncapped = false;
    };

    // takes ownership of a handle from the C ABI; `part` marks a loaded share
    static csr5 adopt(dim_csr5* h, const part_t* part = nullptr) {
        csr5 m{h};
        if (part) {
            m.part_ = *part;
            m.is_part_ = true;
        }
        return m;
    }

    // csr5::from_csr (csr5.h:791-803)
    static csr5 from_csr(const csr& m, const csr5_options& o = {}) {
        if (m.row_start_offsets.size() != size_t{m.dimensions.rows} + 1 || m.values.size() != m.col_indices.size())
            throw std::invalid_argument{"csr arrays do not match the dimensions"};
        return from_csr(m.dimensions.rows, m.dimensions.cols, m.values.size(), m.row_start_offsets.data(),
                        m.col_indices.data(), m.values.data(), o);
    }
    static csr5 from_csr(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const double* vals, const csr5_options& o = {}) {
        dim_csr5_opts opts{o.sigma, o.keep_f32 ? DIM_CSR5_KEEP_F32 : 0u};
        dim_csr5* h = nullptr;
        check(dim_csr5_from_csr(rows, cols, nnz, row_ptr, col_idx, vals, &opts, &h));
        return csr5{h};
    }

    calibrator_t create_calibrator() const { return {}; }

    // y = A x over device spans; x.size() must equal the column count (the
    // reference throws invalid_argument otherwise, csr5.h:820-821). The
    // partial strategy is for a part from h5::load_csr5_partial: y then holds
    // its rows first_row_idx() .. last_row_idx() (csr5.h:585-594).
    template<spmv_strategy S = spmv_strategy::absolute>
    void spmv(::dim::gpu::spmv_data_t<double> d, void* stream = nullptr) const {
        if (S == spmv_strategy::partial && !is_part_)
            throw std::invalid_argument{"spmv<partial> needs a part loaded by h5::load_csr5_partial "
                                        "(or spmv_block on a whole matrix)"};
        check(dim_csr5_spmv(h_.get(), DIM_F64, d.x.data(), d.x.size(), d.y.data(), d.y.size(), stream));
    }
    template<spmv_strategy S = spmv_strategy::absolute>
    void spmv(::dim::gpu::spmv_data_t<float> d, void* stream = nullptr) const {
        if (S == spmv_strategy::partial && !is_part_)
            throw std::invalid_argument{"spmv<partial> needs a part loaded by h5::load_csr5_partial"};
        check(dim_csr5_spmv(h_.get(), DIM_F32, d.x.data(), d.x.size(), d.y.data(), d.y.size(), stream));
    }

    // The same product with x and y in HOST memory, as the reference's
    // callers hold them (cache_aligned_vector, benchmarks/csr5/dim/main.cpp:
    // 58-68): copied to the device and back, returns once y is written.
    void spmv_host(::dim::gpu::spmv_data_t<double> d, void* stream = nullptr) const {
        check(dim_csr5_spmv_host(h_.get(), DIM_F64, d.x.data(), d.x.size(), d.y.data(), d.y.size(), stream));
    }
    void spmv_host(::dim::gpu::spmv_data_t<float> d, void* stream = nullptr) const {
        check(dim_csr5_spmv_host(h_.get(), DIM_F32, d.x.data(), d.x.size(), d.y.data(), d.y.size(), stream));
    }

    // spmv<partial> on tiles [tile_begin, tile_end): y_block has
    // block_rows(...) entries starting at row first_row_idx
    struct block_t {
        uint32_t first_row, n_rows;
    };
    block_t block_rows(uint32_t tile_begin, uint32_t tile_end) const {
        block_t b{};
        check(dim_csr5_block_rows(h_.get(), tile_begin, tile_end, &b.first_row, &b.n_rows));
        return b;
    }
    void spmv_block(uint32_t tile_begin, uint32_t tile_end, std::span<const double> x, std::span<double> y_block,
                    void* stream = nullptr) const {
        check(dim_csr5_spmv_block(h_.get(), DIM_F64, tile_begin, tile_end, x.data(), x.size(), y_block.data(),
                                  y_block.size(), stream));
    }

    // Y = A X for k right-hand sides (row-major, device)
    void spmm(uint32_t k, const double* X, double* Y, void* stream = nullptr) const {
        check(dim_csr5_spmm(h_.get(), DIM_F64, k, X, Y, stream));
    }
    // the same on column-major blocks with leading dimensions ldx, ldy
    void spmm_colmajor(uint32_t k, const double* X, uint64_t ldx, double* Y, uint64_t ldy,
                       void* stream = nullptr) const {
        check(dim_csr5_spmm_colmajor(h_.get(), DIM_F64, k, X, ldx, Y, ldy, stream));
    }

    dimensions_t dimensions() const {
        const auto i = info();
        return {i.rows, i.cols};
    }
    uint64_t non_zero_count() const { return info().nnz; }
    uint32_t tile_count() const { return info().tile_count; }
    uint32_t tail_partition_start() const { return info().tail_partition_start; }
    // the first and last row a product writes (csr5.h:849-855): the whole
    // matrix, or a loaded part's rows
    uint32_t first_row_idx() const { return is_part_ ? part_.first_row : 0u; }
    uint32_t last_row_idx() const {
        return is_part_ ? part_.first_row + part_.n_rows - 1 : dimensions().rows - 1;
    }
    uint64_t val_offset() const { return is_part_ ? part_.val_offset : 0u; }
    // csr5.h:845-847: the first element does not start a row (only a loaded
    // part can begin inside a row)
    bool first_tile_uncapped() const { return is_part_ && part_.first_tile_uncapped; }
    bool skip_tail() const { return is_part_ && part_.skip_tail; }
    uint32_t sigma() const { return info().sigma; }
    static constexpr uint32_t omega() { return 64; }
    dim_csr5_info info() const {
        dim_csr5_info i{};
        check(dim_csr5_get_info(h_.get(), &i));
        return i;
    }

    dim_csr5* handle() const { return h_.get(); }

  private:
    struct deleter {
        void operator()(dim_csr5* p) const { (void)dim_csr5_destroy(p); }
    };
    explicit csr5(dim_csr5* h) : h_(h) {}
    std::unique_ptr<dim_csr5, deleter> h_;
    part_t part_{};
    bool is_part_ = false;
};

struct cg_options {
    double rtol = 1e-8;
    uint32_t max_iters = 5000;
    bool f32_refined = false;
};

// conjugate_gradient (thesis.tex:2133-2160) on device vectors; x holds x0.
// Returns the result; a non-SPD matrix throws (runtime_error), reaching
// max_iters does not (check result.converged).
inline dim_cg_result conjugate_gradient(const csr5& A, const double* b, double* x, const cg_options& o = {},
                                        void* stream = nullptr) {
    dim_cg_opts opts{};
    opts.rtol = o.rtol;
    opts.max_iters = o.max_iters;
    opts.precision = o.f32_refined ? DIM_CG_F32_REFINED : DIM_CG_F64;
    dim_cg_result res{};
    const dim_status st = dim_cg(A.handle(), b, x, &opts, &res, stream);
    if (st != DIM_NOT_CONVERGED)
        check(st);
    return res;
}

// The same solve with b and x in HOST memory (dim_cg_host).
inline dim_cg_result conjugate_gradient_host(const csr5& A, const double* b, double* x, const cg_options& o = {},
                                             void* stream = nullptr) {
    dim_cg_opts opts{};
    opts.rtol = o.rtol;
    opts.max_iters = o.max_iters;
    opts.precision = o.f32_refined ? DIM_CG_F32_REFINED : DIM_CG_F64;
    dim_cg_result res{};
    const dim_status st = dim_cg_host(A.handle(), b, x, &opts, &res, stream);
    if (st != DIM_NOT_CONVERGED)
        check(st);
    return res;
}

} // namespace dim::gpu

// The reference's spelling of these types (csr5.h:254-260, csr.h:19-23,
// base.h:10-14).
namespace dim::mat {
using ::dim::gpu::dimensions_t;
template<typename V = double, std::size_t Sigma = 16, std::size_t Omega = 4, typename U = std::uint32_t,
         template<typename> class Storage = std::vector>
using csr5 = ::dim::gpu::csr5;
template<typename V = double, template<typename> class Storage = std::vector>
using csr = ::dim::gpu::csr;
template<typename V = double>
using calibrator_t = ::dim::gpu::calibrator_t;
} // namespace dim::mat

#endif // DIM_GPU_CSR5_HPP
