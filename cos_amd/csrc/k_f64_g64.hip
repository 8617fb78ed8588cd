// Explicit instantiation of the FP64 per-pixel kernel for L <= 64 (64-lane groups,
// grid-stride; used for the plain FP64 run and the fallback of wide apertures).
#include "capon_fused_impl.h"

namespace cosk {
template hipError_t launch_one<double, 64, 1, 0, true, true, false>(const CaponLaunch&, int, hipStream_t);
}  // namespace cosk
