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
// v_mfma_f32_32x32x16_bf16 layouts (wave64): lane l supplies A[i = l%32][k = 8(l/32) + q]
// and B[k = 8(l/32) + q][j = l%32], q < 8, and holds D[i][j] for j = l%32,
// i = 8(v/4) + 4(l/32) + v%4 in accumulator element v < 16.  For one MFMA step s the lane
// therefore needs the four complex values A[n][l%32], n = 8s + 4(l/32) + 0..3, for P and
// Q alike: each lane reads its own subarray elements and nothing is exchanged between
// lanes.  The split uses the hardware conversion (v_cvt_pk_bf16_f32, round to nearest
// even); the imaginary part's P is the real part's operands with (Re, Im) -> (Im, -Re),
// a swap and a sign flip of the BF16 bits.
#include "cos_device.h"
#include "capon_kernels.h"

namespace cosk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

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

  // offset of subarray n = t K + k in the staged rows, t MP + k; past the last subarray
  // (up to a whole number of steps) the zero tail of row 0 (offset M, + lane < MP)
  int* const offs = reinterpret_cast<int*>(rows + TW * MP);
  const int NS = (N + 7) / 8;   // MFMA steps of 8 subarrays (16 real inner indices)
  for (int n = lane; n < 8 * NS; n += 64) {
    const int t = n / K;
    offs[n] = n < N ? t * MP + (n - t * K) : M;
  }
  __syncthreads();

  const int c = lane & 31, h = lane >> 5;
  // four independent accumulators: hi*hi and the two cross products, real and imaginary
  f32x16 srh = {}, srx = {}, sih = {}, six = {};
  // the same trip count for both half-waves: every MFMA runs with the full wave
  for (int s8 = 0; s8 < NS; ++s8) {
    float v[8];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 e = rows[offs[8 * s8 + 4 * h + q] + c];
      v[2 * q] = e.x;
      v[2 * q + 1] = e.y;
    }
    bf16x8 qh, ql;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      qh[q] = (__bf16)v[q];
      ql[q] = (__bf16)(v[q] - (float)qh[q]);
    }
    // (Re, Im) -> (Im, -Re) on the BF16 bits
    const u16x8 hb = __builtin_bit_cast(u16x8, qh), lb = __builtin_bit_cast(u16x8, ql);
    u16x8 ph, pl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ph[2 * q] = hb[2 * q + 1];
      ph[2 * q + 1] = hb[2 * q] ^ 0x8000;
      pl[2 * q] = lb[2 * q + 1];
      pl[2 * q + 1] = lb[2 * q] ^ 0x8000;
    }
    const bf16x8 pih = __builtin_bit_cast(bf16x8, ph), pil = __builtin_bit_cast(bf16x8, pl);
    srh = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qh, qh, srh, 0, 0, 0);
    sih = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, qh, sih, 0, 0, 0);
    srx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(qh, ql, srx, 0, 0, 0);
    six = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, ql, six, 0, 0, 0);
    srx = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ql, qh, srx, 0, 0, 0);
    six = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pil, qh, six, 0, 0, 0);
  }
  const f32x16 sr = srh + srx, si = sih + six;

  // tr(S) = sum_t sum_m |x_t[m]|^2 #{(k, l): k + l = m, k < K, l < L}, from the data (FP32)
  float trp = 0.f;
  for (int t = 0; t < TW; ++t)
    for (int m = lane; m < M; m += 64) {
      const float2 e = rows[t * MP + m];
      const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
      trp = fmaf((float)cnt, fmaf(e.x, e.x, e.y * e.y), trp);
    }
  const float delta = d * Grp<64>::sum(trp) / (float)L;

  float2* Rp = R + (size_t)p * L * L;
#pragma unroll
  for (int vv = 0; vv < 16; ++vv) {
    const int i = 8 * (vv / 4) + 4 * h + (vv % 4);
    if (i < L && c < L) Rp[i * L + c] = make_float2(sr[vv] + (i == c ? delta : 0.f), -si[vv]);
  }
}

hipError_t launch_cov_mfma(const CaponLaunch& a, hipStream_t s) {
  if (a.rowL || a.plist || a.L < 1 || a.L > 32 || a.M < a.L || !a.R) return hipErrorInvalidValue;
  const int MP = ((a.M + 32 + 3) / 4) * 4;
  const int N = (2 * a.Y + 1) * (a.M - a.L + 1);
  const size_t lds = (size_t)(2 * a.Y + 1) * MP * sizeof(float2) + (size_t)((N + 7) / 8 * 8) * sizeof(int);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const int npix = a.Nx * (a.Nz - 2 * a.Y);
  if (npix == 0) return hipSuccess;
  hipLaunchKernelGGL(cov_mfma_kernel, dim3(npix), dim3(64), lds, s, a.x, a.Nx, a.Nz, a.Mstride, a.M, a.L, a.off,
                     a.Y, a.d, MP, a.R);
  return hipGetLastError();
}

}  // namespace cosk
