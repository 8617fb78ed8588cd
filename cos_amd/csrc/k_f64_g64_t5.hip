// Explicit instantiation of the fused Capon kernel in FP64 for subarrays up to 64, windows
// of up to five rows (the recursion's alpha and beta in registers): the launches over the
// fallback pixel list (one block per entry, and the bounded persistent grid).
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<double, 64, 1, 5, false, false, false>(const CaponLaunch&, int, hipStream_t);
template hipError_t launch_one<double, 64, 1, 5, false, true, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
