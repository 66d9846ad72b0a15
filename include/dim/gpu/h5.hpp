// dim/gpu/h5.hpp - the reference's HDF5 loading calls (dim::io::h5,
// include/dim/io/h5.h:107-127) returning GPU matrices, over dim_h5.h. Header
// only; link libdim_h5.so and libdim_gpu.so.
//
//   reference                                        here
//   h5::load_csr5(path, group)                       the same: a dim::gpu::csr5 (either group layout)
//   h5::load_csr5_partial(group, {idx, total_count}) h5::load_csr5_partial(path, group, {idx, total_count}):
//                                                    one rank's share, to multiply with spmv<partial>
//
// Errors throw std::invalid_argument for bad arguments and std::runtime_error
// otherwise (the reference's h5 errors are runtime_error, h5/err.h:8-16).
#ifndef DIM_GPU_H5_HPP
#define DIM_GPU_H5_HPP

#include "../../dim_h5.h"
#include "csr5.hpp"

#include <cstddef>
#include <filesystem>
#include <stdexcept>
#include <string>

namespace dim::io::h5 {

// include/dim/io/h5.h:119-123
struct partial_identifier_t {
    std::size_t idx;
    std::size_t total_count;
};

inline void check_h5(dim_status st) {
    if (st == DIM_SUCCESS)
        return;
    const std::string msg = std::string(dim_status_string(st)) + ": " + dim_h5_last_error();
    if (st == DIM_INVALID_ARGUMENT)
        throw std::invalid_argument{msg};
    throw std::runtime_error{msg};
}

inline ::dim::gpu::csr5 load_csr5(const std::filesystem::path& path, const std::string& group_name,
                                  const ::dim::gpu::csr5_options& o = {}) {
    dim_csr5_opts opts{o.sigma, o.keep_f32 ? DIM_CSR5_KEEP_F32 : 0u};
    dim_csr5* h = nullptr;
    check_h5(dim_csr5_load_h5(path.c_str(), group_name.c_str(), &opts, &h));
    return ::dim::gpu::csr5::adopt(h);
}

inline ::dim::gpu::csr5 load_csr5_partial(const std::filesystem::path& path, const std::string& group_name,
                                          partial_identifier_t part, const ::dim::gpu::csr5_options& o = {}) {
    dim_h5_csr5_part p{};
    check_h5(dim_h5_load_csr5_partial(path.c_str(), group_name.c_str(), static_cast<uint32_t>(part.idx),
                                      static_cast<uint32_t>(part.total_count), &p));
    dim_csr5_opts opts{o.sigma, o.keep_f32 ? DIM_CSR5_KEEP_F32 : 0u};
    dim_csr5* h = nullptr;
    const dim_status st = dim_csr5_from_csr(p.block.rows, p.block.cols, p.block.nnz, p.block.row_ptr,
                                            p.block.col_idx, p.block.vals, &opts, &h);
    const ::dim::gpu::csr5::part_t info{p.first_row, p.block.rows, p.val_offset, p.skip_tail != 0,
                                        p.first_tile_uncapped != 0};
    dim_h5_free_csr5_part(&p);
    ::dim::gpu::check(st);
    return ::dim::gpu::csr5::adopt(h, &info);
}

} // namespace dim::io::h5

#endif // DIM_GPU_H5_HPP
