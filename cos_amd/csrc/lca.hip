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
// With a 1 x 1 filter this is the per-pixel argmin of |Z| (lcacx.pyx:119-125).  This is
// the centred mode: the reference's general branch indexes its |Z|^2 array as [w][y][x]
// although it fills it as [w][x][y] (lcacx.pyx:186-206), which only agrees for square
// images, and its all-ones range branch never writes row Ym; the centred kernels use the
// intended [x][y] neighbourhood at every interior pixel.
//
// Reference mode (lca_ref_* kernels) reproduces what lcacx.pyx computes, index for index,
// in its three branches: the 1 x 1 argmin of |Z| (:122-127), the all-ones range filter
// by FP64 sliding sums in its order (:139-182: rows 1+Ym .. Ny-Ym-1, row Y_AVG written
// with window 0 by the first segment and overwritten unless it lies past the sliding
// rows, the rest 0), and any other filter through the swapped flat index
// w Nx Ny + (y+yy-Ym) Ny + (x+xx-Xm) into the |Z|^2 array filled as [w][x][y]
// (:185-217).  Their FP64 arithmetic is uncontracted, as the reference's x86-64 build
// (no FMA) rounds every product and sum on its own.
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

// ---- reference mode ----

// |z|^2 in FP64 (both products are exact for float inputs; the sum is rounded once)
__device__ __forceinline__ double sq_f64(float2 v) {
#pragma clang fp contract(off)
  const double re = v.x, im = v.y;
  return re * re + im * im;
}

// s + a b with the product rounded before the sum
__device__ __forceinline__ double add_prod_f64(double s, double a, double b) {
#pragma clang fp contract(off)
  return s + a * b;
}

// a - b, rounded on its own (the difference of lcacx.pyx:171's s + (a - b))
__device__ __forceinline__ double diff_f64(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}

// 1 x 1 filter with a nonzero weight: the first window of smallest |Z| at every pixel.
__global__ __launch_bounds__(256) void lca_ref_argmin_kernel(const LcaLaunch a) {
  const long pix = (long)blockIdx.x * 256 + threadIdx.x;
  if (pix >= (long)a.Nx * a.Ny) return;
  const float2* zp = a.Z + pix * a.Nw;
  double best = sq_f64(zp[0]);
  int bw = 0;
  for (int w = 1; w < a.Nw; ++w) {
    const double v = sq_f64(zp[w]);
    if (v < best) {
      best = v;
      bw = w;
    }
  }
  if (a.img) a.img[pix] = zp[bw];
  if (a.win) a.win[pix] = bw;
}

// All-ones range filter: one block of 256 threads per beam x; thread t < Nw keeps the
// running sum of window t in FP64, in the reference's order: s <- s + (a - b) row by row,
// the difference rounded first (lcacx.pyx:168-171).  Per chunk of sliding rows the block
// forms every row's differences |Z(y+Ym)|^2 - |Z(y-Ym-1)|^2 in parallel into LDS
// [rows][Nw] (each thread's loads all in flight before the first is used), the window
// threads add them in row order (32 rows' differences read ahead) and write the sums back
// in place, and the block takes the rows for the first minimum over w, the gathers of the
// chosen Z again all in flight.  (One wave per beam that waited on each row's loads took
// 0.53 ms for 256 x 2,044 rows.)
constexpr int LCA_REF_THREADS = 256;
__global__ __launch_bounds__(LCA_REF_THREADS) void lca_ref_sliding_kernel(const LcaLaunch a) {
  extern __shared__ double sb[];
  constexpr int NT = LCA_REF_THREADS;
  const int x = blockIdx.x, t = threadIdx.x;
  const int YA = a.wy, Ym = YA / 2, Ny = a.Ny, Nw = a.Nw, CH = a.rows_per_chunk;
  const float2* Zx = a.Z + (size_t)x * Ny * Nw;
  const size_t ob = (size_t)x * Ny;
  const int ylo = 1 + Ym, yhi = Ny - Ym;   // rows of the sliding loop
  // rows it does not write: 0 and window 0, except row Y_AVG when it lies past them,
  // which keeps the first segment's window 0 (Y_AVG >= ylo always)
  for (int y = t; y < Ny; y += NT) {
    if (y >= ylo && y < yhi) continue;
    const float2 v = y == YA ? Zx[(size_t)YA * Nw] : make_float2(0.f, 0.f);
    if (a.img) a.img[ob + y] = v;
    if (a.win) a.win[ob + y] = 0;
  }
  double s = 0.0;   // window t's running sum (t < Nw)
  if (t < Nw)
    for (int y = 0; y < YA; ++y) s = s + sq_f64(Zx[(size_t)y * Nw + t]);
  for (int y0 = ylo; y0 < yhi; y0 += CH) {
    const int nr = min(CH, yhi - y0), ne = nr * Nw;
    for (int e0 = t; e0 < ne; e0 += NT * 16) {
      float2 za[16], zb[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int e = e0 + NT * u, r = e / Nw, w = e - r * Nw, y = y0 + r;
        za[u] = e < ne ? Zx[(size_t)(y + Ym) * Nw + w] : make_float2(0.f, 0.f);
        zb[u] = e < ne ? Zx[(size_t)(y - Ym - 1) * Nw + w] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (e0 + NT * u < ne) sb[e0 + NT * u] = diff_f64(sq_f64(za[u]), sq_f64(zb[u]));
    }
    __syncthreads();
    if (t < Nw) {
      double acc = s;
      for (int r0 = 0; r0 < nr; r0 += 32) {
        double d[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) d[u] = r0 + u < nr ? sb[(r0 + u) * Nw + t] : 0.0;
#pragma unroll
        for (int u = 0; u < 32; ++u)
          if (r0 + u < nr) {
            acc = acc + d[u];
            sb[(r0 + u) * Nw + t] = acc;
          }
      }
      s = acc;
    }
    __syncthreads();
    for (int r0 = t; r0 < nr; r0 += NT * 4) {
      int bw[4];
      float2 zv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + NT * u;
        bw[u] = 0;
        if (r < nr) {
          const double* row = sb + r * Nw;
          double best = row[0];
          for (int w = 1; w < Nw; ++w)
            if (row[w] < best) {
              best = row[w];
              bw[u] = w;
            }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + NT * u;
        zv[u] = r < nr ? Zx[(size_t)(y0 + r) * Nw + bw[u]] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + NT * u;
        if (r < nr) {
          if (a.img) a.img[ob + y0 + r] = zv[u];
          if (a.win) a.win[ob + y0 + r] = bw[u];
        }
      }
    }
    __syncthreads();
  }
}

// Any other filter: one thread per pixel; the |Z|^2 entry at flat index f of the
// reference's [Nw][Nx][Ny] array is that of Z at pixel f mod (Nx Ny), window f / (Nx Ny).
__global__ __launch_bounds__(256) void lca_ref_general_kernel(const LcaLaunch a) {
  const long pix = (long)blockIdx.x * 256 + threadIdx.x;
  if (pix >= (long)a.Nx * a.Ny) return;
  const int x = (int)(pix / a.Ny), y = (int)(pix - (long)x * a.Ny);
  const int Ym = a.wy / 2, Xm = a.wx / 2;
  const long NP = (long)a.Nx * a.Ny;
  float2 best_v = make_float2(0.f, 0.f);
  int bw = 0;
  if (y >= Ym && y < a.Ny - Ym && x >= Xm && x < a.Nx - Xm) {
    double best = 0.0;
    for (int w = 0; w < a.Nw; ++w) {
      double sw = 0.0;
      for (int yy = 0; yy < a.wy; ++yy)
        for (int xx = 0; xx < a.wx; ++xx) {
          const long f = (long)w * NP + (long)(y + yy - Ym) * a.Ny + (x + xx - Xm);
          const long wq = f / NP, r = f - wq * NP;
          sw = add_prod_f64(sw, sq_f64(a.Z[r * a.Nw + wq]), (double)a.wrf[yy * a.wx + xx]);
        }
      if (w == 0 || sw < best) {
        best = sw;
        bw = w;
      }
    }
    best_v = a.Z[pix * a.Nw + bw];
  }
  if (a.img) a.img[pix] = best_v;
  if (a.win) a.win[pix] = bw;
}

}  // namespace

hipError_t launch_lca(const LcaLaunch& a, hipStream_t s) {
  const long npx = (long)a.Nx * a.Ny;
  if (npx <= 0) return hipSuccess;
  if (a.branch != 0) {
    const long blocks = (npx + 255) / 256;
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    if (a.branch == 1) {
      hipLaunchKernelGGL(lca_ref_argmin_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    } else if (a.branch == 2) {
      if (a.Nw > LCA_REF_THREADS || a.rows_per_chunk < 1 || a.Ny <= a.wy) return hipErrorInvalidValue;
      const size_t lds = (size_t)a.rows_per_chunk * a.Nw * sizeof(double);
      if (lds > 64 * 1024) return hipErrorInvalidValue;
      hipLaunchKernelGGL(lca_ref_sliding_kernel, dim3(a.Nx), dim3(LCA_REF_THREADS), lds, s, a);
    } else if (a.branch == 3) {
      hipLaunchKernelGGL(lca_ref_general_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (a.Nw <= LCA_NW_MAX && a.wy / 2 <= LCA_YM_MAX && a.Nx <= 65535) return launch_tile_nw<1>(a, s);
  const long blocks = (npx + 255) / 256;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(lca_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace cosk
