// Internal (non-ABI) launch interface shared by the kernel and API translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cosk {

// Everything one launch of the per-pixel Capon kernel needs.  Sizes and pointers
// are plain; all pointers are device pointers.  Pixel p = ix*(Nz-2Y) + j is the
// reference's output order (RTC/capon_kernel.cu:364-391); its centre row is c = j+Y.
struct CaponLaunch {
  const float2* x;     // cube [Nx][Nz][Mstride] (complex float interleaved)
  int Nx, Nz, Mstride, Y;
  int L, M, off;       // uniform subarray length / active elements / first element
  int Lcap, Mcap;      // bounds the host promised for every row (L <= Lcap, M <= Mcap)
  float d;             // uniform relative diagonal loading
  const int* rowL;     // optional per-centre-row tables, length Nz (override L/M/off/d)
  const int* rowM;
  const int* rowOff;
  const float* rowD;
  const int* plist;    // optional pixel list (values p); null = all pixels
  const int* pcount;   // optional device count for plist (overrides npix when set)
  int pbase;           // with pcount: list entries before plist; the launch covers entries
                       // pbase .. *pcount-1 of the whole list, at most npix of them
  int npix;            // number of pixels (list capacity when pcount is set; grid bound with rlist)
  int jlo, jn;         // without plist: output rows jlo .. jlo+jn-1 of every beam (npix = Nx*jn)
  const int* rlist;    // optional device list of output rows (with rcount, its device length):
  const int* rcount;   // the launch covers rows rlist[0 .. *rcount-1] of every beam, pixel
                       // g -> beam g / *rcount, row rlist[g % *rcount]; npix bounds Nx * *rcount
  float* P;            // [npix_total] power 1/(a^H R_L^-1 a), R normalised by K(2Y+1)
  float2* z;           // [npix_total] amplitude w^H g
  float2* w;           // [npix_total][wstride] weights (first L entries)
  int wstride;
  float2* R;           // [npix_total][Rstride] loaded unnormalised S_L, conj-stored
  int Rstride;
  uint8_t* flags;      // [npix_total] bit0: cond bound > cond_thresh, bit1: non-PD pivot
  float cond_thresh;
  int refine;          // FP32 path: one step of iterative refinement of y
  int cov_only;        // stop after the (optional) R output: BuildR::getR
  // adjoint (backward launch only): upstream dl/dP [npix], accumulated dl/dx (cube layout,
  // atomics; caller zeroes it), per-pixel dl/dd [npix]
  const float* gP;
  float2* gx;
  float* gd;
  // steering (optional): nsteer sines; outputs P/z are [p][nsteer], w is [p][nsteer][wstride]
  const float* steer_sin;
  int nsteer;
  // with steering, angles per block (0: all): the grid's second dimension takes the chunks,
  // each block factorising its pixels and solving for angles blockIdx.y * steer_chunk ..
  int steer_chunk;
  // options: forward-backward averaging, power-Capon amplitude, unloaded power
  int fb, zmode, pmode;
  // optional loaded covariance per pixel in the reference R layout (mem[r*L+c] =
  // conj(S_L[r][c]) for the pixel's L, pixel stride Rstride): the runtime-window FP32
  // kernels then solve with it instead of building the covariance
  const float2* Rin;
  // matrix-core covariance work inside the 16-lane, two-column FP32 kernel (windows of up
  // to five rows): 1 per-snapshot Grams (no R output); 0 and 2 the recursion's rank-two
  // corrections summed per window where a wave qualifies; 3 the recursion on the VALU only
  int mfma_gram;
};

// Row lists of a table-driven frame, built on the device: output row j (0 <= j < nzp) goes
// to list bucket[L(j + Y)] (L clamped to 0 .. 64), in increasing j; lists [nb][nzp] and
// counts [nb].  One block; nb <= 8.
struct RowBuckets {
  unsigned char b[65];
};
hipError_t plan_rows(const int* rowL, int nzp, int Y, RowBuckets bk, int nb, int* lists, int* counts, hipStream_t s);

// Loaded covariance of the pixels of output rows jlo .. jlo+jn-1 (all beams) on the matrix
// cores (BF16 hi/lo split, FP32 accumulation) into a.R + p*Rs in the reference layout;
// ntl 32 x 32 tiles per side (L <= 32 ntl), per-row tables as in CaponLaunch.
hipError_t launch_cov_mfma(const CaponLaunch& a, int ntl, int Mmax, int Rs, hipStream_t s);

// Lane shapes of the FP32 fused kernel, the autotuner's choices per subarray length:
// 1 = 16 lanes x 1 column (L <= 16), 2 = 16 lanes x 2 columns (L <= 32), 3 = 64 lanes
// (L <= 64), 4 = 16 lanes x 3 columns (L <= 48).  default_shape is
// the 16-lane shape with the fewest columns that holds L (four pixels per wave beat the
// 64-lane shape's one wherever a 16-lane group holds L), else the 64-lane shape.
constexpr int kShapes = 4;
inline int shape_cap(int shape) {
  return shape == 1 ? 16 : shape == 2 ? 32 : shape == 3 ? 64 : shape == 4 ? 48 : 0;
}
inline bool shape_fits(int shape, int L) { return shape_cap(shape) >= L; }
inline int default_shape(int L) { return L <= 16 ? 1 : L <= 32 ? 2 : L <= 48 ? 4 : 3; }

// Returns hipSuccess or an error; `why` receives a static message on argument errors.
hipError_t launch_capon_f32(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why);
hipError_t launch_capon_f32_shape(const CaponLaunch& a, int shape, int Mmax, hipStream_t s, const char** why);
hipError_t launch_capon_f64(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why);
// The FP64 kernel over a device-side pixel list (a.plist, a.pcount, a.pbase): one block per
// group of list entries over a grid sized by a.npix (persist = false; entries past the
// count exit at once), or a bounded persistent grid (persist = true).
hipError_t launch_capon_f64_list(const CaponLaunch& a, int Lmax, int Mmax, bool persist, hipStream_t s,
                                 const char** why);
hipError_t launch_capon_bwd(const CaponLaunch& a, int Lmax, int Mmax, hipStream_t s, const char** why);
// Uniform L in (64, 72] (or up to 128 with cov_only): one block per pixel in LDS (wide.hip);
// the default options only (no tables, lists, steering).
hipError_t launch_capon_wide(const CaponLaunch& a, hipStream_t s);

}  // namespace cosk
