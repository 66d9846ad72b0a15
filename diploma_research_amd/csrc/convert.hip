// CSR -> CSR5 (omega = 64) conversion on the GPU, from device CSR arrays.
//
// Produces exactly the format of the host builder (format.cpp), following
// the reference conversion (include/dim/mat/storage_formats/csr5.h:
// generate_tile_pointer 411-431, set_bit_flag 323-340, generate_tile_offset
// 363-397, transpose 463-493) with data-parallel steps:
//   1. tile_ptr by binary search of each tile's first element in row_ptr;
//   2. one thread per tile column: the bits of the rows that start in it;
//   3. dirty tiles from a prefix count of empty rows;
//   4. one wavefront per tile: segment count and tile flags;
//   5. segment -> row maps of dirty tiles (scan of their segment counts);
//   6. the per-tile transpose of values and column indices;
//   7. the list of empty rows (stream compaction).
#include "handle.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <memory>

namespace dimgpu {

namespace {

constexpr int kT = 256;

__device__ __forceinline__ uint32_t upper_bound_idx(const uint32_t* a, uint32_t n, uint64_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (static_cast<uint64_t>(a[mid]) <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo - 1;
}

// first index with a[i] >= v
__device__ __forceinline__ uint32_t lower_bound_idx(const uint32_t* a, uint32_t n, uint64_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (static_cast<uint64_t>(a[mid]) < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// row_ptr[0] == 0, monotone, row_ptr[rows] == nnz, col_idx < cols
__global__ void k_validate(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rp, const uint32_t* ci,
                           int* bad) {
    const size_t tid = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t r = tid; r < rows; r += stride)
        if (rp[r + 1] < rp[r])
            atomicOr(bad, 1);
    for (size_t k = tid; k < nnz; k += stride)
        if (ci[k] >= cols)
            atomicOr(bad, 2);
    if (tid == 0 && (rp[0] != 0 || rp[rows] != nnz))
        atomicOr(bad, 4);
}

__global__ void k_tile_ptr(uint32_t T, uint32_t rows, uint64_t nnz, uint64_t tile, const uint32_t* rp,
                           uint32_t* tile_ptr) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > T)
        return;
    if (t == T) {
        tile_ptr[T] = rows;
        return;
    }
    const uint64_t b = t * tile < nnz ? t * tile : nnz;
    tile_ptr[t] = upper_bound_idx(rp, rows + 1, b);
}

// empty[r] = 1 if row r has no element (r < rows), 0 for r == rows
__global__ void k_empty_flags(uint32_t rows, const uint32_t* rp, uint32_t* empty) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows)
        empty[r] = rp[r] == rp[r + 1] ? 1u : 0u;
    else if (r == rows)
        empty[r] = 0u;
}

template<typename W>
__global__ void k_bit_flags(uint32_t T, uint32_t rows, uint64_t nnz, uint32_t sigma, const uint32_t* rp, W* bits) {
    const size_t w = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (w >= static_cast<size_t>(T) * kOmega)
        return;
    const uint64_t lo = w * sigma;
    const uint64_t hi = lo + sigma < nnz ? lo + sigma : nnz;
    uint32_t b = 0;
    for (uint32_t r = lower_bound_idx(rp, rows + 1, lo); r < rows && rp[r] < hi; ++r)
        if (rp[r + 1] > rp[r])
            b |= 1u << (rp[r] - lo);
    bits[w] = static_cast<W>(b);
}

// one wavefront per tile: dirty bit, segment count, tile flags
template<typename W>
__global__ void k_tile_info(uint32_t T, uint32_t rows, uint32_t* tile_ptr, const uint32_t* empty_prefix,
                            const W* bits, uint8_t* tile_flags, uint32_t* nseg_dirty) {
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= T)
        return;
    const uint32_t b = static_cast<uint32_t>(bits[static_cast<size_t>(t) * kOmega + lane]);
    int n = __popc(b | (lane == 0 ? 1u : 0u));
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        n += __shfl_xor(n, d, 64);
    if (lane == 0) {
        const uint32_t start = tile_ptr[t];
        const uint32_t stop = tile_ptr[t + 1] < rows ? tile_ptr[t + 1] : rows;
        const bool dirty = stop > start && empty_prefix[stop] > empty_prefix[start];
        uint8_t fl = 0;
        if (b & 1u)
            fl |= kStartsRow;
        if (n > 1)
            fl |= kMultiSeg;
        tile_flags[t] = fl;
        nseg_dirty[t] = dirty ? static_cast<uint32_t>(n) : 0u;
    }
}

template<typename W>
__global__ void k_tile_offsets(uint32_t T, uint32_t rows, uint32_t sigma, const uint32_t* rp, uint32_t* tile_ptr,
                               const W* bits, const uint32_t* nseg_dirty, const uint32_t* tdo_ptr, uint32_t* tdo) {
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= T || nseg_dirty[t] == 0)
        return; // wave-uniform
    const uint32_t b = static_cast<uint32_t>(bits[static_cast<size_t>(t) * kOmega + lane]) | (lane == 0 ? 1u : 0u);
    const int nb = __popc(b);
    int incl = nb;
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(incl, d, 64);
        if (lane >= d)
            incl += u;
    }
    const uint32_t start = tile_ptr[t];
    uint32_t k = tdo_ptr[t] + static_cast<uint32_t>(incl - nb);
    const uint64_t tile = uint64_t{kOmega} * sigma;
    for (uint32_t r = 0; r < sigma; ++r) {
        if (!((b >> r) & 1u))
            continue;
        const uint64_t g = t * tile + static_cast<uint64_t>(lane) * sigma + r;
        tdo[k++] = upper_bound_idx(rp, rows + 1, g) - start;
    }
    // every lane has read tile_ptr[t] before lane 0 marks it dirty
    __builtin_amdgcn_wave_barrier();
    if (lane == 0)
        tile_ptr[t] = start | kDirty;
}

__global__ void k_transpose(uint64_t padded, uint64_t nnz, uint32_t sigma, const uint32_t* ci, const double* v,
                            uint32_t* ci_out, double* v_out, float* v32_out) {
    const uint64_t tile = uint64_t{kOmega} * sigma;
    for (size_t dst = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; dst < padded;
         dst += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint64_t t = dst / tile, rem = dst % tile;
        const uint64_t r = rem / kOmega, c = rem % kOmega;
        const uint64_t src = t * tile + c * sigma + r;
        const bool in = src < nnz;
        const double a = in ? v[src] : 0.0;
        ci_out[dst] = in ? ci[src] : 0u;
        v_out[dst] = a;
        if (v32_out)
            v32_out[dst] = static_cast<float>(a);
    }
}

__global__ void k_row_index(uint32_t rows, uint32_t* idx) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows)
        idx[r] = r;
}

template<typename W>
dim_status convert(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* rp, const uint32_t* ci,
                   const double* v, uint32_t sigma, bool keep_f32, hipStream_t s, dim_csr5** out) {
    std::unique_ptr<dim_csr5, void (*)(dim_csr5*)> h(new dim_csr5, free_handle);
    DIM_HIP_TRY(hipGetDevice(&h->device));
    const uint64_t tile = uint64_t{kOmega} * sigma;
    const uint64_t T64 = (nnz + tile - 1) / tile;
    if (T64 >= kDirty)
        return fail(DIM_INVALID_ARGUMENT, "too many tiles");
    const uint32_t T = static_cast<uint32_t>(T64);
    DevFormat& d = h->f;
    d.rows = rows;
    d.cols = cols;
    d.sigma = sigma;
    d.tiles = T;
    d.nnz = nnz;
    d.padded = T64 * tile;

    // temporaries
    std::vector<void*> tmp;
    struct Tmp {
        std::vector<void*>& v;
        ~Tmp() {
            for (void* p : v)
                (void)hipFree(p);
        }
    } tmp_guard{tmp};
    auto talloc = [&](void** p, size_t bytes) {
        hipError_t e = hipMalloc(p, bytes ? bytes : 1);
        if (e == hipSuccess)
            tmp.push_back(*p);
        return e;
    };

    int* bad = nullptr;
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&bad), sizeof(int)));
    DIM_HIP_TRY(hipMemsetAsync(bad, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_validate, dim3(1024), dim3(kT), 0, s, rows, cols, nnz, rp, ci, bad);
    int hbad = 0;
    DIM_HIP_TRY(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    DIM_HIP_TRY(hipStreamSynchronize(s));
    if (hbad & 1)
        return fail(DIM_INVALID_ARGUMENT, "row_ptr must be non-decreasing");
    if (hbad & 2)
        return fail(DIM_INVALID_ARGUMENT, "column index out of range");
    if (hbad & 4)
        return fail(DIM_INVALID_ARGUMENT, "row_ptr must start at 0 and end at nnz");

    DIM_HIP_TRY(dev_alloc(h.get(), d.tile_ptr, size_t{T} + 1));
    hipLaunchKernelGGL(k_tile_ptr, dim3((T + kT) / kT), dim3(kT), 0, s, T, rows, nnz, tile, rp, d.tile_ptr);

    // empty rows: flags, prefix count, compacted list
    uint32_t *empty = nullptr, *empty_prefix = nullptr, *row_ids = nullptr, *row_sel = nullptr, *nsel = nullptr;
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&row_sel), (size_t{rows} + 1) * 4));
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&empty), (size_t{rows} + 1) * 4));
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&empty_prefix), (size_t{rows} + 1) * 4));
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&row_ids), (size_t{rows} + 1) * 4));
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&nsel), 4));
    hipLaunchKernelGGL(k_empty_flags, dim3((rows + kT) / kT), dim3(kT), 0, s, rows, rp, empty);
    hipLaunchKernelGGL(k_row_index, dim3((rows + kT) / kT), dim3(kT), 0, s, rows, row_ids);
    size_t tb1 = 0, tb2 = 0, tb3 = 0;
    DIM_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb1, empty, empty_prefix, rows + 1, s));
    DIM_HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, tb2, row_ids, empty, row_sel, nsel, rows, s));
    DIM_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, empty, empty_prefix, size_t{T} + 1, s));
    void* tstore = nullptr;
    const size_t tbytes = std::max(std::max(tb1, tb2), tb3);
    DIM_HIP_TRY(talloc(&tstore, tbytes));
    size_t tb = tbytes;
    DIM_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tstore, tb, empty, empty_prefix, rows + 1, s));

    // bit flags
    W* bits = nullptr;
    DIM_HIP_TRY(dev_alloc(h.get(), bits, std::max<size_t>(size_t{T} * kOmega, 1)));
    d.bit_flag = bits;
    if (T)
        hipLaunchKernelGGL((k_bit_flags<W>), dim3((T * kOmega + kT - 1) / kT), dim3(kT), 0, s, T, rows, nnz, sigma, rp,
                           bits);

    // tile info, dirty segment counts, their scan
    // flag words are read 4 bytes at a time: pad to a multiple of 4 plus 4
    const size_t flag_bytes = (size_t{T} + 1 + 3) / 4 * 4 + 4;
    DIM_HIP_TRY(dev_alloc(h.get(), d.tile_flags, flag_bytes));
    DIM_HIP_TRY(hipMemsetAsync(d.tile_flags, 0, flag_bytes, s));
    uint32_t* nseg_dirty = nullptr;
    DIM_HIP_TRY(talloc(reinterpret_cast<void**>(&nseg_dirty), (size_t{T} + 1) * 4));
    DIM_HIP_TRY(hipMemsetAsync(nseg_dirty, 0, (size_t{T} + 1) * 4, s));
    const uint8_t sentinel = kStartsRow;
    DIM_HIP_TRY(hipMemcpyAsync(d.tile_flags + T, &sentinel, 1, hipMemcpyHostToDevice, s));
    if (T)
        hipLaunchKernelGGL((k_tile_info<W>), dim3((T + 3) / 4), dim3(256), 0, s, T, rows, d.tile_ptr, empty_prefix,
                           static_cast<const W*>(bits), d.tile_flags, nseg_dirty);
    DIM_HIP_TRY(dev_alloc(h.get(), d.tile_desc_offset_ptr, size_t{T} + 1));
    tb = tbytes;
    DIM_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tstore, tb, nseg_dirty, d.tile_desc_offset_ptr, size_t{T} + 1, s));
    uint32_t ntdo = 0;
    DIM_HIP_TRY(hipMemcpyAsync(&ntdo, d.tile_desc_offset_ptr + T, 4, hipMemcpyDeviceToHost, s));
    // empty-row list
    tb = tbytes;
    DIM_HIP_TRY(hipcub::DeviceSelect::Flagged(tstore, tb, row_ids, empty, row_sel, nsel, rows, s));
    uint32_t n_empty = 0;
    DIM_HIP_TRY(hipMemcpyAsync(&n_empty, nsel, 4, hipMemcpyDeviceToHost, s));
    DIM_HIP_TRY(hipStreamSynchronize(s));
    d.any_dirty = ntdo != 0;
    d.n_empty = n_empty;
    DIM_HIP_TRY(dev_alloc(h.get(), d.tile_desc_offset, std::max<size_t>(ntdo, 1)));
    if (ntdo)
        hipLaunchKernelGGL((k_tile_offsets<W>), dim3((T + 3) / 4), dim3(256), 0, s, T, rows, sigma, rp, d.tile_ptr,
                           static_cast<const W*>(bits), static_cast<const uint32_t*>(nseg_dirty),
                           static_cast<const uint32_t*>(d.tile_desc_offset_ptr), d.tile_desc_offset);
    DIM_HIP_TRY(dev_alloc(h.get(), d.empty_rows, std::max<size_t>(n_empty, 1)));
    if (n_empty)
        DIM_HIP_TRY(hipMemcpyAsync(d.empty_rows, row_sel, size_t{n_empty} * 4, hipMemcpyDeviceToDevice, s));

    // transposed values and columns
    DIM_HIP_TRY(dev_alloc(h.get(), d.col_idx, d.padded));
    DIM_HIP_TRY(dev_alloc(h.get(), d.vals64, d.padded));
    if (keep_f32)
        DIM_HIP_TRY(dev_alloc(h.get(), d.vals32, d.padded));
    if (d.padded)
        hipLaunchKernelGGL(k_transpose, dim3(2048), dim3(kT), 0, s, d.padded, nnz, sigma, ci, v, d.col_idx, d.vals64,
                           d.vals32);
    DIM_HIP_TRY(hipGetLastError());

    // scratch (as upload_format)
    Scratch& sc = h->s;
    const size_t Tm = std::max<size_t>(T, 1);
    double *hp = nullptr, *tp = nullptr;
    DIM_HIP_TRY(dev_alloc(h.get(), hp, Tm * kMaxK));
    DIM_HIP_TRY(dev_alloc(h.get(), tp, Tm * kMaxK));
    sc.head = hp;
    sc.tail = tp;
    DIM_HIP_TRY(dev_alloc(h.get(), sc.chunk_flags, Tm));
    DIM_HIP_TRY(dev_alloc(h.get(), sc.dot_chunk, Tm * kMaxK));
    DIM_HIP_TRY(dev_alloc(h.get(), sc.dot_blk, ((Tm + 255) / 256) * kMaxK));
    DIM_HIP_TRY(upload(h.get(), sc.queue, std::vector<uint32_t>(kQueueWords, 0u)));
    DIM_HIP_TRY(hipStreamSynchronize(s));
    // host copy of the tile flags for the launch planner
    std::vector<uint8_t> hflags(size_t{T} + 1);
    DIM_HIP_TRY(hipMemcpy(hflags.data(), d.tile_flags, size_t{T} + 1, hipMemcpyDeviceToHost));
    h->binfo.assign(std::move(hflags));
    d.binfo = &h->binfo;
    // dirty-tile count for get_info
    if (ntdo) {
        std::vector<uint32_t> nsd(T);
        DIM_HIP_TRY(hipMemcpy(nsd.data(), nseg_dirty, size_t{T} * 4, hipMemcpyDeviceToHost));
        for (uint32_t x : nsd)
            h->n_dirty += x ? 1 : 0;
    }
    h->sys_row = {0, rows};
    h->sys_tile = {0, T};
    h->sys_nnz_csr = {0, nnz};
    h->nnz_logical = nnz;
    *out = h.release();
    return DIM_SUCCESS;
}

} // namespace
} // namespace dimgpu

using namespace dimgpu;

extern "C" dim_status dim_csr5_from_csr_device(uint32_t rows, uint32_t cols, uint64_t nnz, const uint32_t* row_ptr,
                                               const uint32_t* col_idx, const double* vals, const dim_csr5_opts* opts,
                                               void* stream, dim_csr5** out) {
    if (!out)
        return fail(DIM_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    if (!row_ptr || (nnz && (!col_idx || !vals)))
        return fail(DIM_INVALID_ARGUMENT, "null CSR array");
    if (rows >= kDirty)
        return fail(DIM_INVALID_ARGUMENT, "row count must be below 2^31");
    if (nnz >= (uint64_t{1} << 32))
        return fail(DIM_INVALID_ARGUMENT, "non-zero count must fit in 32 bits");
    const uint32_t sigma = opts && opts->sigma ? opts->sigma : 16;
    if (!sigma_supported(sigma))
        return fail(DIM_INVALID_ARGUMENT, "unsupported sigma");
    const bool keep_f32 = opts && (opts->flags & DIM_CSR5_KEEP_F32);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (sigma <= 16)
        return convert<uint16_t>(rows, cols, nnz, row_ptr, col_idx, vals, sigma, keep_f32, s, out);
    return convert<uint32_t>(rows, cols, nnz, row_ptr, col_idx, vals, sigma, keep_f32, s, out);
}
