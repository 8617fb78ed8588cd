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
// lca_tile_kernel<NW> (NW <= 16 windows, Ym <= 32): a block owns one beam x and 256
// ranges; for each beam of the filter's x extent it stages |Z|^2 of its 256 + 2 Ym rows
// (all windows, one coalesced pass over contiguous memory) in LDS, and each thread adds
// its neighbourhood from there, so Z is read about (2 Xm + 1) times (once for a filter
// along range) and every |Z|^2 is formed once per staging.  lca_kernel, one thread per
// pixel reading its neighbourhood through the caches, serves the other shapes.
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

constexpr int LCA_T = 256;       // ranges per block of the tiled kernel
constexpr int LCA_NW_MAX = 16;   // windows held per pixel in registers
constexpr int LCA_YM_MAX = 32;   // halo rows

template <int NW>
__global__ __launch_bounds__(LCA_T) void lca_tile_kernel(const LcaLaunch a) {
  extern __shared__ float pw[];                     // [(LCA_T + 2 Ym) * NW] |Z|^2 of one beam
  const int x = blockIdx.y, y0 = blockIdx.x * LCA_T, t = threadIdx.x, y = y0 + t;
  const int Ym = a.wy / 2, Xm = a.wx / 2, rows = LCA_T + 2 * Ym;
  const long pix = (long)x * a.Ny + y;
  // pixels whose neighbourhood leaves the image get 0 and window 0 (a whole block when its
  // beam is within Xm of the edge)
  const bool xin = x >= Xm && x < a.Nx - Xm;
  float s[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) s[w] = 0.f;
  if (xin) {
    for (int xx = 0; xx < a.wx; ++xx) {
      const size_t base = (size_t)(x + xx - Xm) * a.Ny;
      __syncthreads();                                // the previous beam's tile is consumed
      for (int i = t; i < rows * NW; i += LCA_T) {
        const int yr = y0 - Ym + i / NW, w = i - (i / NW) * NW;
        float v = 0.f;
        if (yr >= 0 && yr < a.Ny) {
          const float2 z = a.Z[(base + yr) * NW + w];
          v = fmaf(z.x, z.x, z.y * z.y);
        }
        pw[i] = v;
      }
      __syncthreads();
      for (int yy = 0; yy < a.wy; ++yy) {
        const float wv = a.wrf[yy * a.wx + xx];
        const float* row = pw + (t + yy) * NW;
#pragma unroll
        for (int w = 0; w < NW; ++w) s[w] = fmaf(wv, row[w], s[w]);
      }
    }
  }
  if (y >= a.Ny) return;
  float2 best = make_float2(0.f, 0.f);
  int bw = 0;
  if (xin && y >= Ym && y < a.Ny - Ym) {
    float bs = s[0];
#pragma unroll
    for (int w = 1; w < NW; ++w)
      if (s[w] < bs) {                                // first minimum
        bs = s[w];
        bw = w;
      }
    best = a.Z[pix * NW + bw];
  }
  if (a.img) a.img[pix] = best;
  if (a.win) a.win[pix] = bw;
}

template <int NW>
static hipError_t launch_tile(const LcaLaunch& a, hipStream_t s) {
  const dim3 grid((a.Ny + LCA_T - 1) / LCA_T, a.Nx);
  const size_t lds = (size_t)(LCA_T + 2 * (a.wy / 2)) * NW * sizeof(float);
  hipLaunchKernelGGL(lca_tile_kernel<NW>, grid, dim3(LCA_T), lds, s, a);
  return hipGetLastError();
}

template <int NW>
static hipError_t launch_tile_nw(const LcaLaunch& a, hipStream_t s) {
  if constexpr (NW > LCA_NW_MAX) {
    return hipErrorInvalidValue;
  } else {
    return a.Nw == NW ? launch_tile<NW>(a, s) : launch_tile_nw<NW + 1>(a, s);
  }
}

}  // namespace

hipError_t launch_lca(const LcaLaunch& a, hipStream_t s) {
  const long npx = (long)a.Nx * a.Ny;
  if (npx <= 0) return hipSuccess;
  if (a.Nw <= LCA_NW_MAX && a.wy / 2 <= LCA_YM_MAX && a.Nx <= 65535) return launch_tile_nw<1>(a, s);
  const long blocks = (npx + 255) / 256;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lca_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace cosk
