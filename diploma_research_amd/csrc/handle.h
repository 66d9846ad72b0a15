// The dim_csr5 handle behind the C ABI, and error plumbing shared by the
// C ABI translation units.
#pragma once

#include "dim_internal.h"
#include "kernels.h"
#include "../../include/dim_gpu.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace dimgpu {
struct CgCache;
void free_cg_cache(CgCache* c);
struct BatchCache;
void free_batch_cache(BatchCache* c);
} // namespace dimgpu

struct dim_csr5 {
    dimgpu::DevFormat f;
    dimgpu::Scratch s;
    int device = 0;
    uint32_t n_dirty = 0;
    uint64_t device_bytes = 0;
    std::vector<void*> allocations;
    // block-diagonal batches: system i owns rows [sys_row[i], sys_row[i+1]) and
    // tiles [sys_tile[i], sys_tile[i+1])
    uint32_t n_systems = 1;
    std::vector<uint32_t> sys_row, sys_tile;
    std::vector<uint64_t> sys_nnz_csr;
    // chunks of system s in the chunk table: [sys_chunk[s], sys_chunk[s + 1])
    std::vector<uint32_t> sys_chunk;
    uint32_t* d_sys_chunk = nullptr;
    uint64_t nnz_logical = 0; // stored elements (the format's nnz counts padding inside a batch)
    dimgpu::BoundaryInfo binfo; // host copy of the tile flags (f.binfo points here)
    // Column slices for plain products (capi.cpp): xs is the matrix whose
    // rows are the (column slice s, row i) pairs that hold elements, in the
    // order of s*rows + i. That stacked handle also owns the partial products
    // (ystack, one FP64-sized entry per stacked row), the bit map of the
    // pairs (present, bit s*rows + i) and the set bits before each word of
    // it (present_rank).
    dim_csr5* xs = nullptr;
    uint32_t xs_slices = 0;
    void* ystack = nullptr;
    uint32_t* present = nullptr;
    uint32_t* present_rank = nullptr;
    // column-major staging of row-major multi-RHS blocks (dim_csr5_spmm),
    // grown on demand
    void* mm_scratch = nullptr;
    size_t mm_bytes = 0;
    // device copies of host vectors for the host-pointer entry points
    // (dim_csr5_spmv_host, dim_cg_host): cols + rows FP64-sized entries,
    // allocated on first use
    void* host_stage = nullptr;
    // work buffers, pinned state and captured iteration graph of the last
    // dim_cg / dim_cg_multi solve, kept for the next solve of the same shape
    // (cg.hip); freed with the handle
    dimgpu::CgCache* cg_cache = nullptr;
    // the same for dim_cg_batched (batched.hip): block tables and buffers
    // sized by the batch, kept from the first batched solve
    dimgpu::BatchCache* batch_cache = nullptr;
};

namespace dimgpu {

void set_error(const std::string& msg);
dim_status fail(dim_status st, const std::string& msg);
dim_status hip_fail(hipError_t e, const char* what);

// The staging area of the host-pointer entry points: cols + rows entries of
// 8 bytes on the matrix's device (*x at the start, *y after it), allocated
// with the handle's arrays on first use.
dim_status stage_host_vectors(dim_csr5* A, char** x, char** y);

#define DIM_HIP_TRY(expr)                                                                                          \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return ::dimgpu::hip_fail(e_, #expr);                                                                  \
    } while (0)

// device allocation owned by the handle
template<typename T>
hipError_t dev_alloc(dim_csr5* h, T*& p, size_t count) {
    p = nullptr;
    const size_t bytes = count ? count * sizeof(T) : sizeof(T);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return e;
    h->allocations.push_back(q);
    h->device_bytes += bytes;
    p = static_cast<T*>(q);
    return hipSuccess;
}

template<typename T>
hipError_t upload(dim_csr5* h, T*& p, const std::vector<T>& v) {
    hipError_t e = dev_alloc(h, p, v.size());
    if (e != hipSuccess)
        return e;
    if (!v.empty())
        e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

// Uploads a host format into a fresh handle (allocates scratch too).
// plain_columns: leave out the column codes (products stream col_idx).
dim_status upload_format(const HostFormat& f, bool keep_f32, dim_csr5** out, bool plain_columns = false);

// One system's share of a batched matrix as a stand-alone format view.
DevFormat system_view(const dim_csr5* A, uint32_t s);
void free_handle(dim_csr5* h);

} // namespace dimgpu
