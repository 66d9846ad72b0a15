// dim_stream_read: the HBM read rate this GPU reaches on a plain stream, the
// measured ceiling that the benchmark reports beside the vendor peak. Not on
// the CSR5 path and with no counterpart in the reference.
#include "handle.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace dimgpu {
namespace {

__global__ void k_zero_one(double* p) { *p = 0.0; }

// a 16-byte word (a clang vector type, which the non-temporal load takes)
typedef double dbl2 __attribute__((ext_vector_type(2)));

// Block b reads its own contiguous range [b*per, (b+1)*per) of 16-byte words
// with non-temporal loads, eight in flight per thread, and adds them up so
// that the loads are kept; one atomic add per wavefront. Measured on the
// MI355X, contiguous ranges of four blocks per CU stream faster (6.9 TB/s on
// 4 GiB) than a grid-stride walk (6.1 TB/s).
__global__ __launch_bounds__(256) void k_stream_read(const dbl2* __restrict__ p, size_t n2, size_t per, double* out) {
    constexpr int U = 8;
    const size_t b0 = static_cast<size_t>(blockIdx.x) * per;
    const size_t b1 = b0 + per < n2 ? b0 + per : n2;
    double acc = 0.0;
    size_t i = b0 + threadIdx.x;
    for (; i + (U - 1) * 256 < b1; i += U * 256) {
        dbl2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            v[u] = __builtin_nontemporal_load(p + i + u * 256);
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc += v[u].x + v[u].y;
    }
    for (; i < b1; i += 256) {
        const dbl2 a = __builtin_nontemporal_load(p + i);
        acc += a.x + a.y;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        acc += __shfl_xor(acc, d, 64);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(out, acc);
}

} // namespace
} // namespace dimgpu

extern "C" dim_status dim_stream_read(const void* data, uint64_t bytes, double* sum, void* stream) {
    using namespace dimgpu;
    if (!sum || (bytes && !data) || (bytes % 16) || (reinterpret_cast<uintptr_t>(data) % 16))
        return fail(DIM_INVALID_ARGUMENT, "dim_stream_read needs a 16-byte aligned buffer of a multiple of 16 bytes");
    int dev = 0, cus = 0;
    DIM_HIP_TRY(hipGetDevice(&dev));
    DIM_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const auto st = static_cast<hipStream_t>(stream);
    const size_t n2 = bytes / 16;
    const uint32_t grid = static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>(size_t(cus) * 4, n2 / 256)));
    const size_t per = (n2 + grid - 1) / grid;
    hipLaunchKernelGGL(k_zero_one, dim3(1), dim3(1), 0, st, sum);
    if (n2)
        hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, st, static_cast<const dbl2*>(data), n2, per, sum);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DIM_SUCCESS : hip_fail(e, "dim_stream_read");
}
