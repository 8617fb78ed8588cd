// Explicit instantiation of the adjoint (backward) Capon kernel, 32 columns, M <= 64,
// windows of up to five rows, the forward recomputation's rank-two corrections on the
// matrix cores (FP16 hi/lo split).
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<float, 16, 2, 5, false, false, 64, false, 2>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
