// Time-domain interpolation beamformer (TDIB): the delay stage in front of Capon.
//
// Semantics (framework/beamformer/delay/tdibC.pyx:9-67, paths relative to the reference
// root): for beam nx and pixel ny at (xarr[nx][ny], yarr[nx][ny]), with the origin at
// xx = mean(x_r)/2, zz = mean(D)/2,
//   r_t = sqrt((xx + xarr)^2 + yarr^2 + zz^2)                    transmitter -> pixel
//   r_r = sqrt((xx + xarr - x_r[n])^2 + yarr^2 + (zz - D[n])^2)   pixel -> hydrophone n
//   t   = (r_t + r_r) / c
//   v_n = interp(mfdata[n], t) * exp(j 2 pi fc t)                  linear interpolation on
//         the time axis; outside the axis the value is the fill (the reference's
//         interp1d(bounds_error=False) gives NaN).  On a general (strictly increasing)
//         axis the interval is scipy's: hi = the first index with taxis[hi] >= t
//         (searchsorted, left), clipped to [1, n_t - 1], lo = hi - 1, by binary search.
//   image[ny][nx] = (1/n_h) sum_n v_n
// The delayed channels v_n are also written as the cube [n_x][n_y][n_h], which is the
// layout the Capon kernels read ([beam][range][element]).
//
// A block of four waves owns 64 consecutive range pixels of one beam; lanes run over the
// pixels and the waves over the hydrophones (interleaved), so for one hydrophone the lanes
// read neighbouring samples of its channel (neighbouring pixels have neighbouring delays:
// coalesced reads).  The delayed values of a chunk of 64 hydrophones go through an LDS
// tile [pixel][hydrophone] (rows padded against bank conflicts) and leave as the cube's
// contiguous [pixel][hydrophone] rows with coalesced stores; the image sums meet in LDS.
// Geometry, interpolation index and the phase's reduction to [0, 2 pi) are evaluated in
// FP64: t ~ 0.1 s and fc ~ 1e5 Hz put the phase near 1e4 rad, where FP32 rounding of t
// alone would cost ~1e-2 rad.  The interpolation weight, the interpolation, the sine and
// cosine of the reduced phase and the rotation are FP32 (relative error ~3e-7 against the
// FP64 formula), the image sum too.
#include "cos_device.h"
#include "aux_stages.h"

namespace cosk {

namespace {

constexpr int TD_TP = 64;          // pixels per block (one per lane)
constexpr int TD_HC = 64;          // hydrophones per LDS tile
constexpr int TD_PAD = TD_HC + 1;  // tile row stride in float2

__global__ __launch_bounds__(256) void tdib_kernel(const TdibLaunch a, const int tiles_per_beam) {
  __shared__ float2 tile[TD_TP * TD_PAD];
  __shared__ float2 part[4][TD_TP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nx = blockIdx.x / tiles_per_beam, ny0 = (blockIdx.x - nx * tiles_per_beam) * TD_TP;
  const int ny = ny0 + lane;
  const bool valid = ny < a.n_y;
  const long pix = (long)nx * a.n_y + (valid ? ny : ny0);
  // origin: xx = mean(x_r)/2, zz = mean(D)/2 (a wave sum; n_h is small)
  double sx = 0.0, sd = 0.0;
  for (int n = lane; n < a.n_h; n += 64) {
    sx += a.x_r[n];
    sd += a.D[n];
  }
  const double xx = Grp<64>::sum(sx) / (2.0 * a.n_h), zz = Grp<64>::sum(sd) / (2.0 * a.n_h);
  const double inv_dt = 1.0 / a.dt;
  const double xa = a.xarr[pix], ya = a.yarr[pix];
  const double ex = xx + xa, ya2 = ya * ya;
  const double r_t = sqrt(ex * ex + ya2 + zz * zz);
  float2 acc = make_float2(0.f, 0.f);
  for (int n0 = 0; n0 < a.n_h; n0 += TD_HC) {
    const int nc = min(TD_HC, a.n_h - n0);
    for (int j = w; j < nc; j += 4) {
      const int n = n0 + j;
      const double dx = ex - a.x_r[n], dz = zz - a.D[n];
      const double r_r = sqrt(dx * dx + ya2 + dz * dz);
      const double t = (r_t + r_r) * a.inv_c;
      // linear interpolation: interval i .. i+1 and weight f of t in it
      float2 v = make_float2(a.fill, a.fill);
      int i = 0;
      float f = 0.f;
      bool in = false;
      if (a.taxis) {
        const double* ta = a.taxis;
        in = t >= ta[0] && t <= ta[a.n_t - 1];
        if (in && a.n_t > 1) {
          int lo = 0, hi = a.n_t - 1;            // ta[lo] <= t <= ta[hi]: first index >= t
          while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ta[mid] < t) lo = mid;
            else hi = mid;
          }
          if (ta[lo] >= t) hi = lo;              // t == ta[0]
          hi = hi < 1 ? 1 : hi;
          i = hi - 1;
          f = (float)((t - ta[i]) / (ta[hi] - ta[i]));
        }
      } else {
        const double u = (t - a.t0) * inv_dt;
        in = u >= 0.0 && u <= (double)(a.n_t - 1);
        if (in) {
          i = (int)u;                            // 0 <= i <= n_t - 1
          if (i > a.n_t - 2) i = a.n_t > 1 ? a.n_t - 2 : 0;
          f = (float)(u - (double)i);            // 1 at the last sample
        }
      }
      if (in) {
        const float2 s0 = a.mf[(size_t)n * a.n_t + i];
        const float2 s1 = a.mf[(size_t)n * a.n_t + (a.n_t > 1 ? i + 1 : i)];
        const float re = fmaf(f, s1.x - s0.x, s0.x);
        const float im = fmaf(f, s1.y - s0.y, s0.y);
        // mix to carrier: exp(j 2 pi fc t), the phase reduced in FP64 before sincospi
        double ph = 2.0 * a.fc * t;              // in units of pi
        ph -= 2.0 * floor(0.5 * ph);             // [0, 2)
        float sn, cs;
        sincospif((float)ph, &sn, &cs);
        v = make_float2(fmaf(re, cs, -im * sn), fmaf(re, sn, im * cs));
      }
      acc.x += v.x;
      acc.y += v.y;
      tile[lane * TD_PAD + j] = v;
    }
    __syncthreads();
    if (a.cube) {
      // the tile's rows are contiguous in the cube: [pixel][hydrophone n0 .. n0+nc-1]
      float2* dst = a.cube + ((size_t)nx * a.n_y + ny0) * a.n_h + n0;
      for (int e = threadIdx.x; e < TD_TP * nc; e += 256) {
        const int pp = e / nc, jj = e - pp * nc;
        if (ny0 + pp < a.n_y) dst[(size_t)pp * a.n_h + jj] = tile[pp * TD_PAD + jj];
      }
    }
    __syncthreads();
  }
  if (a.image) {
    part[w][lane] = acc;
    __syncthreads();
    if (w == 0 && valid) {
      const float2 s1 = part[1][lane], s2 = part[2][lane], s3 = part[3][lane];
      const float re = (acc.x + s1.x) + (s2.x + s3.x), im = (acc.y + s1.y) + (s2.y + s3.y);
      a.image[(size_t)ny * a.n_x + nx] = make_float2(re / a.n_h, im / a.n_h);
    }
  }
}

}  // namespace

hipError_t launch_tdib(const TdibLaunch& a, hipStream_t s) {
  if ((long)a.n_x * a.n_y <= 0 || a.n_h <= 0) return hipSuccess;
  const int tiles_per_beam = (a.n_y + TD_TP - 1) / TD_TP;
  const long blocks = (long)a.n_x * tiles_per_beam;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tdib_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, tiles_per_beam);
  return hipGetLastError();
}

}  // namespace cosk
