// Explicit instantiations of the adjoint (backward) Capon kernel, 32 columns, M <= 64.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 2, 5, false, false, 64>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 2, 0, false, false, 64>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
