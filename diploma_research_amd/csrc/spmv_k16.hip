// Multi-RHS CSR5 product kernels for right-hand-side counts up to 16.
#include "spmv_impl.h"

namespace dimgpu {
namespace detail {

uint32_t launch_pair_k16(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, bool dot,
                        hipStream_t st) {
    return f32 ? launch_pair<float, 16>(A, tp, fp, dot, st, nullptr) : launch_pair<double, 16>(A, tp, fp, dot, st, nullptr);
}

} // namespace detail
} // namespace dimgpu
