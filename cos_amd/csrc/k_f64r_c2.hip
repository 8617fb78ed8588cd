// Explicit instantiations of the FP64 per-pixel kernel with the covariance by the
// displacement recursion, 32 columns: the launches over the fallback pixel list (one block
// per group of entries, and the bounded persistent grid).
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<double, 16, 2, 0, false, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<double, 16, 2, 0, false, true, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
