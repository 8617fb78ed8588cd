// Explicit instantiation of the fused Capon kernel, 64-lane groups, window bound 9.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 64, 1, 9, false, false, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
