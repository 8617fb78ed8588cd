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
//         interp1d(bounds_error=False) gives NaN)
//   image[ny][nx] = (1/n_h) sum_n v_n
// The delayed channels v_n are also written as the cube [n_x][n_y][n_h], which is the
// layout the Capon kernels read ([beam][range][element]).
//
// One wave per run of TD_PPW consecutive pixels (the origin sums are formed once per
// wave), lanes over hydrophones (the cube row of a pixel is written with one coalesced
// store per 64 hydrophones).  Geometry, interpolation index and the phase's reduction to
// [0, 2 pi) are evaluated in FP64: t ~ 0.1 s and fc ~ 1e5 Hz put the phase near 1e4 rad,
// where FP32 rounding of t alone would cost ~1e-2 rad.  The interpolation weight, the
// interpolation, the sine and cosine of the reduced phase and the rotation are FP32
// (relative error ~3e-7 against the FP64 formula), the image sum of 64 terms too.
#include "cos_device.h"
#include "aux_stages.h"

namespace cosk {

namespace {

constexpr int TD_PPW = 8;   // consecutive pixels per wave

__global__ __launch_bounds__(256) void tdib_kernel(const TdibLaunch a) {
  const int lane = threadIdx.x % 64;
  const long npx = (long)a.n_x * a.n_y;
  const long p0 = ((long)blockIdx.x * 4 + threadIdx.x / 64) * TD_PPW;
  if (p0 >= npx) return;
  // origin: xx = mean(x_r)/2, zz = mean(D)/2 (a wave sum; n_h is small)
  double sx = 0.0, sd = 0.0;
  for (int n = lane; n < a.n_h; n += 64) {
    sx += a.x_r[n];
    sd += a.D[n];
  }
  const double xx = Grp<64>::sum(sx) / (2.0 * a.n_h), zz = Grp<64>::sum(sd) / (2.0 * a.n_h);
  const double inv_dt = 1.0 / a.dt;
  for (int k = 0; k < TD_PPW; ++k) {
    const long pix = p0 + k;
    if (pix >= npx) break;
    const int nx = (int)(pix / a.n_y), ny = (int)(pix - (long)nx * a.n_y);
    const double xa = a.xarr[pix], ya = a.yarr[pix];
    const double ex = xx + xa;
    const double r_t = sqrt(ex * ex + ya * ya + zz * zz);
    float2 acc = make_float2(0.f, 0.f);
    for (int n = lane; n < a.n_h; n += 64) {
      const double dx = ex - a.x_r[n], dz = zz - a.D[n];
      const double r_r = sqrt(dx * dx + ya * ya + dz * dz);
      const double t = (r_t + r_r) * a.inv_c;
      // linear interpolation on the uniform axis t0 + i dt, i < n_t
      const double u = (t - a.t0) * inv_dt;
      float2 v = make_float2(a.fill, a.fill);
      if (u >= 0.0 && u <= (double)(a.n_t - 1)) {
        int i = (int)u;                          // 0 <= i <= n_t - 1
        if (i > a.n_t - 2) i = a.n_t > 1 ? a.n_t - 2 : 0;
        const float f = (float)(u - (double)i);  // 1 at the last sample
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
      if (a.cube) a.cube[((size_t)nx * a.n_y + ny) * a.n_h + n] = v;
    }
    if (a.image) {
      acc.x = Grp<64>::sum(acc.x);
      acc.y = Grp<64>::sum(acc.y);
      if (lane == 0) a.image[(size_t)ny * a.n_x + nx] = make_float2(acc.x / a.n_h, acc.y / a.n_h);
    }
  }
}

}  // namespace

hipError_t launch_tdib(const TdibLaunch& a, hipStream_t s) {
  const long npx = (long)a.n_x * a.n_y;
  if (npx <= 0) return hipSuccess;
  const long blocks = (npx + 4 * TD_PPW - 1) / (4 * TD_PPW);
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tdib_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace cosk
