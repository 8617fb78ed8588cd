// Internal launch interfaces of the stages either side of the Capon kernels: the TDIB
// delay stage that produces the cube and the LCA window selection.
#pragma once
#include <hip/hip_runtime.h>

namespace cosk {

struct TdibLaunch {
  const float2* mf;        // [n_h][n_t] matched-filtered channel data
  int n_h, n_t;
  double t0, dt;           // uniform time axis t0 + i dt (when taxis is null)
  const double* taxis;     // optional [n_t] strictly increasing time axis
  const double* x_r;       // [n_h] hydrophone along-track positions
  const double* D;         // [n_h] hydrophone depths
  double inv_c, fc;        // 1/sound speed, carrier frequency
  const double* xarr;      // [n_x][n_y] pixel x
  const double* yarr;      // [n_x][n_y] pixel y
  int n_x, n_y;
  float fill;              // value outside the time axis
  float2* cube;            // [n_x][n_y][n_h] delayed channels (optional)
  float2* image;           // [n_y][n_x] delay-and-sum image (optional)
};

struct LcaLaunch {
  const float2* Z;         // [Nx][Ny][Nw] images, one per window
  int Nx, Ny, Nw;
  const float* wrf;        // [wy][wx] neighbourhood weights (odd sizes)
  int wy, wx;
  float2* img;             // [Nx][Ny] selected value (optional)
  int* win;                // [Nx][Ny] selected window (optional)
  int branch;              // 0: centred neighbourhood; the reference's own branches:
                           // 1 argmin |Z| (1 x 1 filter), 2 all-ones range filter
                           // (sliding FP64 sums), 3 any other filter (swapped flat index)
  int rows_per_chunk;      // branch 2: sliding rows whose sums are staged in LDS at once
};

hipError_t launch_tdib(const TdibLaunch& a, hipStream_t s);
hipError_t launch_lca(const LcaLaunch& a, hipStream_t s);

}  // namespace cosk
