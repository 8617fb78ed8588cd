// Explicit instantiation of the fused Capon kernel with one refinement step, 16 lanes x 3
// columns, windows of up to five rows, rows and columns 0..31 of the recursion on the matrix
// cores (FP16 hi/lo split), the rest on the VALU.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 3, 5, false, false, 0, true, 2>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
