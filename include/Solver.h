/*
 * Solver.h -- header-only C++ shim with the reference's class and signature
 * (framework/beamformer/capon/RealTimeCapon/Solver.h:17-44, ISolver.h:16-29), forwarding
 * to cos_solver_solve in cos_capon.h (column-major A, 1 <= N <= 72), so
 * RealTimeCapon/Solver.pxd and Solver.pyx bind unchanged.
 */
#pragma once
#include <Complex.h>
#include "cos_capon.h"

class Solver {
public:
   Solver(void) {}
   ~Solver(void) {}

   int solve(Complex<float>* &x, Complex<float>* &A, Complex<float>* &b, int &N, int &batch,
             bool &x_on_gpu, bool &A_on_gpu, bool &b_on_gpu)
   {
      static_assert(sizeof(Complex<float>) == sizeof(cos_cf32), "Complex<float> must be {float re, im}");
      return cos_solver_solve(reinterpret_cast<cos_cf32*>(x), reinterpret_cast<const cos_cf32*>(A),
                              reinterpret_cast<const cos_cf32*>(b), N, batch, x_on_gpu ? 1 : 0,
                              A_on_gpu ? 1 : 0, b_on_gpu ? 1 : 0);
   }
};
