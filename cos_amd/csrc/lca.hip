// Low-complexity adaptive (LCA) beamformer: per pixel, pick the window whose output
// power, summed over a weighted neighbourhood, is lowest.
//
// Semantics (framework/beamformer/lca/lcacx.pyx:27-217, paths relative to the reference
// root): Z[x][y][w] holds the image beamformed with window w.  For every pixel whose
// (2 Ym + 1) x (2 Xm + 1) neighbourhood lies inside the image,
//   s_w = sum_{yy, xx} wrf[yy][xx] |Z[x + xx - Xm][y + yy - Ym][w]|^2,
//   w*  = the first w with the smallest s_w,
//   img[x][y] = Z[x][y][w*], win[x][y] = w*;
// pixels nearer the edge get 0 (as in the reference, whose output arrays start at zero).
// With a 1 x 1 filter this is the per-pixel argmin of |Z| (lcacx.pyx:119-125).  The
// reference's general branch indexes its |Z|^2 array as [w][y][x] although it fills it as
// [w][x][y] (lcacx.pyx:186-206), which only agrees for square images; this kernel uses the
// intended [x][y] neighbourhood.
//
// One thread per pixel; the neighbourhood reads of adjacent threads (adjacent y) overlap
// and come from L1/L2, so the kernel streams Z about once.
#include "cos_device.h"
#include "aux_stages.h"

namespace cosk {

namespace {

__global__ __launch_bounds__(256) void lca_kernel(const LcaLaunch a) {
  const long pix = (long)blockIdx.x * 256 + threadIdx.x;
  if (pix >= (long)a.Nx * a.Ny) return;
  const int x = (int)(pix / a.Ny), y = (int)(pix - (long)x * a.Ny);
  const int Ym = a.wy / 2, Xm = a.wx / 2;
  float2 best = make_float2(0.f, 0.f);
  int bw = 0;
  if (y >= Ym && y < a.Ny - Ym && x >= Xm && x < a.Nx - Xm) {
    float bs = 0.f;
    for (int w = 0; w < a.Nw; ++w) {
      float sw = 0.f;
      for (int yy = 0; yy < a.wy; ++yy)
        for (int xx = 0; xx < a.wx; ++xx) {
          const float2 v = a.Z[(((size_t)(x + xx - Xm) * a.Ny) + (y + yy - Ym)) * a.Nw + w];
          sw = fmaf(a.wrf[yy * a.wx + xx], fmaf(v.x, v.x, v.y * v.y), sw);
        }
      if (w == 0 || sw < bs) {
        bs = sw;
        bw = w;
      }
    }
    best = a.Z[(size_t)pix * a.Nw + bw];
  }
  if (a.img) a.img[pix] = best;
  if (a.win) a.win[pix] = bw;
}

}  // namespace

hipError_t launch_lca(const LcaLaunch& a, hipStream_t s) {
  const long npx = (long)a.Nx * a.Ny;
  if (npx <= 0) return hipSuccess;
  const long blocks = (npx + 255) / 256;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lca_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace cosk
