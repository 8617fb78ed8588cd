// Explicit instantiations of the fused Capon kernel (one translation unit per
// variant group so the library builds in parallel): 16 lanes x 3 columns, L <= 48.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 3, 5, false, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 3, 9, false, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 3, 0, false, false, 0>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
