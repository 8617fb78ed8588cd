// Explicit instantiation of the fused Capon kernel, 16 lanes x 3 columns, windows of up to
// five rows, the recursion's rank-two corrections for rows and columns 0..31 on the matrix
// cores (FP16 hi/lo split), the rest on the VALU.
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 3, 5, false, false, 0, false, 2>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
