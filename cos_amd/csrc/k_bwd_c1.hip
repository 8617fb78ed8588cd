// Explicit instantiations of the adjoint (backward) Capon kernel, 16 columns.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 1, 5, false, false, 64>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 1, 5, false, false, 128>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 1, 0, false, false, 64>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<float, 16, 1, 0, false, false, 128>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
