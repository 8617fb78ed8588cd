// Explicit instantiation of the fused Capon kernel with one refinement step, 16 lanes x 2
// columns, windows of up to five rows, the recursion's rank-two corrections on the matrix
// cores (FP16 hi/lo split), for launches whose rows are padded past 128 elements.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 2, 5, false, false, 0, true, 3>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
