// Explicit instantiation of the fused Capon kernel in FP64 for subarrays up to 64:
// 64-lane groups, covariance by the displacement recursion, runtime window (with the
// reference options), over a pixel range or a device-side pixel list.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<double, 64, 1, 0, false, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<double, 64, 1, 0, false, true, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
