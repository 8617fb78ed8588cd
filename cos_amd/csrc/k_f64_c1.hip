// Explicit instantiations of the FP64 per-pixel kernel (direct covariance), 16 columns:
// full-frame launch and the bounded-grid launch over the fallback pixel list.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<double, 16, 1, 0, true, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<double, 16, 1, 0, true, true, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
