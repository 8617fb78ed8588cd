// Explicit instantiations of the fused Capon kernel (one translation unit per
// variant group so the library builds in parallel).
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 2, 0, false, false, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
