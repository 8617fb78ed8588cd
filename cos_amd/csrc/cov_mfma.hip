// Covariance of every pixel on the matrix cores: the complex rank-k update
// S = sum_t sum_k u_{t,k} u_{t,k}^H, u_{t,k}[l] = x_t[k+l], as two real 32 x 32 products
// with the inner dimension running over the window's subarrays, each operand split into
// a BF16 high part and a BF16 low part (x = hi + lo to 16 significant bits) and three
// products hi*hi + hi*lo + lo*hi accumulated in FP32.
//
// Semantics as the reference's buildR_kernel (framework/beamformer/capon/RealTimeCapon/
// buildR_kernel.cu:45-161): unnormalised S, loading d*tr(S)/L on the diagonal, stored
// conjugated and column-major, mem[r*L + c] = conj(S_L[r][c]).
//
// Real form.  With A[n][r] = x_t[k+r] (n enumerates (t, k)), S = A^T conj(A):
//   Re S[i][j] = sum_n (Re A[n][i] Re A[n][j] + Im A[n][i] Im A[n][j])
//   Im S[i][j] = sum_n (Im A[n][i] Re A[n][j] - Re A[n][i] Im A[n][j]).
// The inner index n' = 2n + part interleaves real and imaginary parts, so both are
// products P Q with the same Q[n'][j] = (Re, Im)(A[n][j]) and P[i][n'] = Q[n'][i] for the
// real part, P[i][n'] = (Im, -Re)(A[n][i]) for the imaginary part.
//
// v_mfma_f32_32x32x8_bf16 layouts (wave64): lane l supplies A[i = l%32][k = 4(l/32) + q]
// and B[k = 4(l/32) + q][j = l%32], q < 4, and holds D[i][j] for j = l%32,
// i = 8(v/4) + 4(l/32) + v%4 in accumulator element v < 16.  For one MFMA step s the lane
// therefore needs the two complex values A[n][l%32] at n = 4s + 2(l/32) and n + 1, for P
// and Q alike: each lane reads its own subarray element twice per step and nothing is
// exchanged between lanes.
#include "cos_device.h"
#include "capon_kernels.h"

namespace cosk {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (inputs are finite)
__device__ __forceinline__ unsigned bf16_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

struct HiLo {
  bf16x4 hi, lo;
};
__device__ __forceinline__ HiLo split4(float a, float b, float c, float d) {
  const float v[4] = {a, b, c, d};
  HiLo s;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const unsigned h = bf16_bits(v[q]);
    s.hi[q] = (short)h;
    s.lo[q] = (short)bf16_bits(v[q] - __uint_as_float(h << 16));
  }
  return s;
}

// One pixel per wave (block of 64).  The window's 2Y+1 rows are staged in LDS, zero padded
// to MP elements (MP >= M + 32, so every subarray element of a padding row or column reads
// a zero or an in-row value).
__global__ __launch_bounds__(64) void cov_mfma_kernel(const float2* __restrict__ x, int Nx, int Nz, int Mstride,
                                                      int M, int L, int off, int Y, float d, int MP,
                                                      float2* __restrict__ R) {
  extern __shared__ float2 rows[];
  const int TW = 2 * Y + 1, Nzp = Nz - 2 * Y, K = M - L + 1, N = TW * K;
  const int p = blockIdx.x;
  const int ix = p / Nzp, j = p - ix * Nzp;
  const int lane = threadIdx.x;
  const float2* src = x + ((size_t)ix * Nz + j) * Mstride + off;
  for (int t = 0; t < TW; ++t)
    for (int m = lane; m < MP; m += 64) rows[t * MP + m] = m < M ? src[(size_t)t * Mstride + m] : make_float2(0.f, 0.f);
  __syncthreads();

  const int c = lane & 31, h = lane >> 5;
  // A[n][c] = x_t[k + c] with n = t K + k, zero for n >= N
  auto elem = [&](int n) -> float2 {
    if (n >= N) return make_float2(0.f, 0.f);
    const int t = n / K;
    return rows[t * MP + (n - t * K) + c];
  };
  f32x16 sr = {}, si = {};
  // the same trip count for both half-waves: every MFMA runs with the full wave
  const int NS = (N + 3) / 4;
  for (int s4 = 0; s4 < NS; ++s4) {
    const int n0 = 4 * s4 + 2 * h;
    const float2 v0 = elem(n0), v1 = elem(n0 + 1);
    const HiLo q = split4(v0.x, v0.y, v1.x, v1.y);     // Q, and P of the real part
    const HiLo pi = split4(v0.y, -v0.x, v1.y, -v1.x);  // P of the imaginary part
    sr = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(q.hi, q.hi, sr, 0, 0, 0);
    sr = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(q.hi, q.lo, sr, 0, 0, 0);
    sr = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(q.lo, q.hi, sr, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(pi.hi, q.hi, si, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(pi.hi, q.lo, si, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(pi.lo, q.hi, si, 0, 0, 0);
  }

  // tr(S) = sum_t sum_m |x_t[m]|^2 #{(k, l): k + l = m, k < K, l < L}, from the data (FP32)
  float trp = 0.f;
  for (int t = 0; t < TW; ++t)
    for (int m = lane; m < M; m += 64) {
      const float2 v = rows[t * MP + m];
      const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
      trp = fmaf((float)cnt, fmaf(v.x, v.x, v.y * v.y), trp);
    }
  const float delta = d * Grp<64>::sum(trp) / (float)L;

  float2* Rp = R + (size_t)p * L * L;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = 8 * (v / 4) + 4 * h + (v % 4);
    if (i < L && c < L) Rp[i * L + c] = make_float2(sr[v] + (i == c ? delta : 0.f), -si[v]);
  }
}

hipError_t launch_cov_mfma(const CaponLaunch& a, hipStream_t s) {
  if (a.rowL || a.plist || a.L < 1 || a.L > 32 || a.M < a.L || !a.R) return hipErrorInvalidValue;
  const int MP = ((a.M + 32 + 3) / 4) * 4;
  const size_t lds = (size_t)(2 * a.Y + 1) * MP * sizeof(float2);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const int npix = a.Nx * (a.Nz - 2 * a.Y);
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(cov_mfma_kernel, dim3(npix), dim3(64), lds, s, a.x, a.Nx, a.Nz, a.Mstride, a.M, a.L, a.off,
                     a.Y, a.d, MP, a.R);
  return hipGetLastError();
}

}  // namespace cosk
