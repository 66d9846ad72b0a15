// Conjugate-gradient solvers (filled in below).
#include "handle.h"

extern "C" {
dim_status dim_cg(const dim_csr5* A, const double*, double*, const dim_cg_opts*, dim_cg_result*, void*) {
    if (!A)
        return dimgpu::fail(DIM_INVALID_ARGUMENT, "matrix handle is null");
    return dimgpu::fail(DIM_UNSUPPORTED, "not built yet");
}
dim_status dim_cg_multi(const dim_csr5*, uint32_t, const double*, double*, const dim_cg_opts*, dim_cg_result*, void*) {
    return dimgpu::fail(DIM_UNSUPPORTED, "not built yet");
}
dim_status dim_csr5_from_csr_batched(uint32_t, const uint32_t*, const uint64_t*, const uint32_t* const*,
                                     const uint32_t* const*, const double* const*, const dim_csr5_opts*, dim_csr5**) {
    return dimgpu::fail(DIM_UNSUPPORTED, "not built yet");
}
dim_status dim_cg_batched(const dim_csr5*, const double*, double*, const dim_cg_opts*, dim_cg_result*, void*) {
    return dimgpu::fail(DIM_UNSUPPORTED, "not built yet");
}
}
