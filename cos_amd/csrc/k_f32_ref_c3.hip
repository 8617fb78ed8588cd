// Explicit instantiation of the fused Capon kernel with one step of iterative
// refinement (runtime window length), one lane shape per translation unit.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 3, 0, false, false, 0, true>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
