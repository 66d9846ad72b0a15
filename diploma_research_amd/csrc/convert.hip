// GPU-side CSR -> CSR5 conversion (filled in below).
#include "handle.h"

extern "C" dim_status dim_csr5_from_csr_device(uint32_t, uint32_t, uint64_t, const uint32_t*, const uint32_t*,
                                               const double*, const dim_csr5_opts*, void*, dim_csr5**) {
    return dimgpu::fail(DIM_UNSUPPORTED, "not built yet");
}
