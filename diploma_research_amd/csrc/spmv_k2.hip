// Multi-RHS CSR5 product kernels for right-hand-side counts up to 2.
#include "spmv_impl.h"

namespace dimgpu {
namespace detail {

void launch_pair_k2(bool f32, const DevFormat& A, const TileParams& tp, const FixParams& fp, uint32_t tgrid,
                    uint32_t fgrid, bool dot, hipStream_t st) {
    if (f32)
        launch_pair<float, 2>(A, tp, fp, tgrid, fgrid, dot, st);
    else
        launch_pair<double, 2>(A, tp, fp, tgrid, fgrid, dot, st);
}

} // namespace detail
} // namespace dimgpu
