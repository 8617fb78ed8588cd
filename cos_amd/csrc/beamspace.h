// Internal launch interface of the sliding-DFT beamspace Capon kernel.
#pragma once
#include <hip/hip_runtime.h>

namespace cosk {

struct BeamspaceLaunch {
  const float2* x;      // cube [Nx][Nz][Mstride]
  int Nx, Nz, Mstride, M, L, off, Y, Nb;
  float d;              // loading relative to tr(R_bs)/Nb
  float rsqrtL, sqrtL;  // L^-1/2 and L^1/2 (host, double-rounded)
  const float2* tw;     // [Nb][L] exp(-j 2 pi beta(b) l / L)
  const float2* omega;  // [Nb]    exp(+j 2 pi beta(b) / L)
  float* P;             // [npx]
  float2* z;            // [npx]
  float2* wbs;          // [npx][Nb] beamspace weights
  float2* w;            // [npx][wstride] element-space weights B^H w_bs
  int wstride;
};

hipError_t launch_beamspace(const BeamspaceLaunch& a, hipStream_t s);

}  // namespace cosk
