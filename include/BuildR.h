/*
 * BuildR.h -- header-only C++ shim with the reference's class and signature
 * (framework/beamformer/capon/RealTimeCapon/BuildR.h:12-32, IBuildR.h:51-64), forwarding
 * to cos_buildr_get_r in cos_capon.h, so RealTimeCapon/BuildR.pxd binds unchanged.
 */
#pragma once
#include <Complex.h>
#include "cos_capon.h"

class BuildR {
public:
   BuildR(void) {}
   ~BuildR(void) {}

   int getR(Complex<float>* &R_in, Complex<float>* &x_in, float &d, int &L, int &Yavg, int &M,
            int &Nx, int &Nz, int &Nb, bool &R_on_gpu, bool &x_on_gpu)
   {
      static_assert(sizeof(Complex<float>) == sizeof(cos_cf32), "Complex<float> must be {float re, im}");
      return cos_buildr_get_r(reinterpret_cast<cos_cf32*>(R_in), reinterpret_cast<const cos_cf32*>(x_in),
                              d, L, Yavg, M, Nx, Nz, Nb, R_on_gpu ? 1 : 0, x_on_gpu ? 1 : 0);
   }
};
