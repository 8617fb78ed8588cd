/*
 * Capon.h -- header-only C++ shim with the reference's class and signature
 * (framework/beamformer/capon/RealTimeCapon/Capon.h:13-125, ICapon.h:20-33), forwarding
 * to the C ABI in cos_capon.h.  With this header first on the include path and
 * libcos_capon.so linked, RealTimeCapon/Capon.pxd and getCaponCUDA.pyx bind unchanged.
 *
 * Complex<float> comes from the caller's Complex.h (framework/lib/Complex.h), whose
 * {re, im} layout is cos_cf32's.
 */
#pragma once
#include <Complex.h>
#include "cos_capon.h"

class Capon {
public:
   Capon(void) {}
   ~Capon(void) {}

   int getCapon(Complex<float>* &z, Complex<float>* &w, Complex<float>* &R, Complex<float>* &x,
                float &d, int &L, int &Yavg, int &M, int &Nx, int &Nz, int &Nb)
   {
      static_assert(sizeof(Complex<float>) == sizeof(cos_cf32), "Complex<float> must be {float re, im}");
      return cos_capon_get_capon(reinterpret_cast<cos_cf32*>(z), reinterpret_cast<cos_cf32*>(w),
                                 reinterpret_cast<cos_cf32*>(R), reinterpret_cast<const cos_cf32*>(x),
                                 d, L, Yavg, M, Nx, Nz, Nb);
   }
};
