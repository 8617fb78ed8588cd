// Variant selection for the fused Capon kernel.
#include "capon_kernels.h"

namespace cosk {

template <typename T, int G, int CPL, int TM, bool DIRECT, bool PERSIST, int BWD, bool REF = false, int MG = 0>
hipError_t launch_one(const CaponLaunch& a, int Mmax, hipStream_t s);

template <int G, int CPL>
static hipError_t launch_f32_tm(const CaponLaunch& a, int Mmax, hipStream_t s) {
  const int TW = 2 * a.Y + 1;
  // one step of iterative refinement: its own kernels, with the fixed window bound when
  // no reference option needs the runtime-window kernels (nor a stored covariance)
  if (a.refine) {
    if (TW <= 5 && !(a.fb || a.zmode || a.pmode || a.Rin)) {
      if constexpr (G == 16 && CPL == 2) {   // the recursion's rank-two terms on the matrix cores
        if (a.mfma_gram == 0 || a.mfma_gram == 2) {
          if (((Mmax + 32 + 4 + 3) / 4) * 4 <= 128) return launch_one<float, 16, 2, 5, false, false, 0, true, 2>(a, Mmax, s);
          return launch_one<float, 16, 2, 5, false, false, 0, true, 3>(a, Mmax, s);   // rows padded past 128
        }
      }
      if constexpr (G == 16 && CPL == 3) {   // rows and columns 0..31 on the matrix cores
        if (a.mfma_gram == 0 || a.mfma_gram == 2) return launch_one<float, 16, 3, 5, false, false, 0, true, 2>(a, Mmax, s);
      }
      return launch_one<float, G, CPL, 5, false, false, 0, true>(a, Mmax, s);
    }
    return launch_one<float, G, CPL, 0, false, false, 0, true>(a, Mmax, s);
  }
  // the optional variants live in the runtime-window kernels
  if (a.fb || a.zmode || a.pmode || a.Rin) return launch_one<float, G, CPL, 0, false, false, 0>(a, Mmax, s);
  if constexpr (G == 16 && CPL == 2) {   // covariance work on the matrix cores inside the kernel
    if (TW <= 5 && a.mfma_gram == 1 && !a.R && !a.cov_only)
      return launch_one<float, 16, 2, 5, false, false, 0, false, 1>(a, Mmax, s);
    if (TW <= 5 && (a.mfma_gram == 0 || a.mfma_gram == 2)) {
      // rows padded past 128 elements: the instantiation that stages them (zero tail)
      if (((Mmax + 32 + 4 + 3) / 4) * 4 > 128) return launch_one<float, 16, 2, 5, false, false, 0, false, 3>(a, Mmax, s);
      return launch_one<float, 16, 2, 5, false, false, 0, false, 2>(a, Mmax, s);
    }
  }
  if constexpr (G == 16 && CPL == 3) {   // rows and columns 0..31 of the recursion on the matrix cores
    if (TW <= 5 && (a.mfma_gram == 0 || a.mfma_gram == 2))
      return launch_one<float, 16, 3, 5, false, false, 0, false, 2>(a, Mmax, s);
  }
  if (TW <= 5) return launch_one<float, G, CPL, 5, false, false, false>(a, Mmax, s);
  if (TW <= 9) return launch_one<float, G, CPL, 9, false, false, false>(a, Mmax, s);
  return launch_one<float, G, CPL, 0, false, false, false>(a, Mmax, s);
}

hipError_t launch_capon_f32_shape(const CaponLaunch& a, int shape, int Mmax, hipStream_t s, const char** why) {
  switch (shape) {
    case 1: return launch_f32_tm<16, 1>(a, Mmax, s);
    case 2: return launch_f32_tm<16, 2>(a, Mmax, s);
    case 3: return launch_f32_tm<64, 1>(a, Mmax, s);
    case 4: return launch_f32_tm<16, 3>(a, Mmax, s);
  }
  if (why) *why = "unknown lane shape";
  return hipErrorInvalidValue;
}

hipError_t launch_capon_f32(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why) {
  if (Lmax <= 64) return launch_capon_f32_shape(a, default_shape(Lmax), Mmax, s, why);
  const hipError_t e = launch_capon_wide(a, s);
  if (e == hipErrorInvalidValue && why) *why = "subarray length beyond the wide kernel (L <= 72; 128 for R only)";
  return e;
}

static hipError_t f64_launch(const CaponLaunch& a, int Lmax, int Mmax, bool list, hipStream_t s, const char** why) {
  // L <= 32: the plain kernel (direct covariance sums, 165 complex MACs per element at
  // M = 64, L = 32, Y = 2) for whole frames, the same-GPU FP64 comparator; over the fallback
  // pixel list the covariance by the FP64 displacement recursion (10 per element)
  if (Lmax <= 16) {
    if (a.plist)
      return list ? launch_one<double, 16, 1, 0, false, true, false>(a, Mmax, s) : launch_one<double, 16, 1, 0, false, false, false>(a, Mmax, s);
    return list ? launch_one<double, 16, 1, 0, true, true, false>(a, Mmax, s) : launch_one<double, 16, 1, 0, true, false, false>(a, Mmax, s);
  }
  if (Lmax <= 32) {
    if (a.plist)
      return list ? launch_one<double, 16, 2, 0, false, true, false>(a, Mmax, s) : launch_one<double, 16, 2, 0, false, false, false>(a, Mmax, s);
    return list ? launch_one<double, 16, 2, 0, true, true, false>(a, Mmax, s) : launch_one<double, 16, 2, 0, true, false, false>(a, Mmax, s);
  }
  if (Lmax <= 64) {   // covariance by the FP64 recursion, the columns in registers
    // the fallback list with windows of up to five rows: the fixed-window kernel (alpha and
    // beta of the recursion held in registers instead of reloaded per window row)
    if (a.plist && 2 * a.Y + 1 <= 5 && !(a.fb || a.zmode || a.pmode))
      return list ? launch_one<double, 64, 1, 5, false, true, false>(a, Mmax, s) : launch_one<double, 64, 1, 5, false, false, false>(a, Mmax, s);
    return list ? launch_one<double, 64, 1, 0, false, true, false>(a, Mmax, s) : launch_one<double, 64, 1, 0, false, false, false>(a, Mmax, s);
  }
  if (why) *why = "L > 64 is not supported by the FP64 kernel";
  return hipErrorInvalidValue;
}

hipError_t launch_capon_f64(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why) {
  return f64_launch(a, Lmax, Mmax, a.pcount != nullptr, s, why);
}

hipError_t launch_capon_f64_list(const CaponLaunch& a, int Lmax, int Mmax, bool persist, hipStream_t s,
                                 const char** why) {
  if (!a.plist || !a.pcount) return hipErrorInvalidValue;
  return f64_launch(a, Lmax, Mmax, persist, s, why);
}

hipError_t launch_capon_bwd(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why) {
  if (a.nsteer > 0) {
    if (why) *why = "the adjoint does not take steering vectors";
    return hipErrorInvalidValue;
  }
  const bool w5 = 2 * a.Y + 1 <= 5;
  if (Mmax > 128) {
    if (why) *why = "the adjoint supports M <= 128";
    return hipErrorInvalidValue;
  }
  // per-lane k/m slots sized for M <= 64 or M <= 128
  if (Lmax <= 16) {
    if (Mmax <= 64)
      return w5 ? launch_one<float, 16, 1, 5, false, false, 64>(a, Mmax, s) : launch_one<float, 16, 1, 0, false, false, 64>(a, Mmax, s);
    return w5 ? launch_one<float, 16, 1, 5, false, false, 128>(a, Mmax, s) : launch_one<float, 16, 1, 0, false, false, 128>(a, Mmax, s);
  }
  if (Lmax <= 32) {
    if (Mmax <= 64) {
      // the forward recomputation's rank-two terms on the matrix cores
      if (w5 && (a.mfma_gram == 0 || a.mfma_gram == 2)) return launch_one<float, 16, 2, 5, false, false, 64, false, 2>(a, Mmax, s);
      return w5 ? launch_one<float, 16, 2, 5, false, false, 64>(a, Mmax, s) : launch_one<float, 16, 2, 0, false, false, 64>(a, Mmax, s);
    }
    return w5 ? launch_one<float, 16, 2, 5, false, false, 128>(a, Mmax, s) : launch_one<float, 16, 2, 0, false, false, 128>(a, Mmax, s);
  }
  if (why) *why = "the adjoint supports L <= 32";
  return hipErrorInvalidValue;
}

}  // namespace cosk
