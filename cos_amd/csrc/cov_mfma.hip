// Covariance of every pixel on the matrix cores: the complex rank-k update
// S = sum_t sum_k u_{t,k} u_{t,k}^H, u_{t,k}[l] = x_t[k+l], as two real products (one
// 32 x 32 tile for L <= 32, the upper three of a 2 x 2 grid of tiles for L <= 64) with the
// inner dimension running over the window's subarrays, each operand split into
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

// One pixel per wave (block of 64), the pixels of one run of output rows (all beams): the
// window's 2Y+1 rows are staged in LDS, zero padded to MP elements (MP >= Mmax + 32 NTL,
// so every subarray element of a padding row or column reads a zero or an in-row value).
// NTL 32 x 32 tiles per side: L <= 32 NTL; for NTL = 2 the three upper blocks are formed
// and the lower one written as the conjugate transpose.  Per-row tables of L, M, first
// element and loading as in the fused kernel.  Output: R + p Rs, the pixel's L x L loaded
// covariance in the reference layout.
template <int NTL>
__global__ __launch_bounds__(64) void cov_mfma_kernel(const CaponLaunch a, const int MP, const int Rs) {
  extern __shared__ float2 rows[];
  const int Y = a.Y, TW = 2 * Y + 1, Nzp = a.Nz - 2 * Y;
  const int b = blockIdx.x;
  int ix, j;
  if (a.rlist) {   // rows rlist[0 .. *rcount-1] of every beam; blocks past them exit
    const int nr = *a.rcount;
    if (b >= a.Nx * nr) return;
    ix = b / nr;
    j = a.rlist[b - ix * nr];
  } else {
    ix = b / a.jn;
    j = a.jlo + (b - ix * a.jn);
  }
  const int p = ix * Nzp + j, crow = j + Y;
  const int L = a.rowL ? a.rowL[crow] : a.L;
  const int M = a.rowM ? a.rowM[crow] : a.M;
  const int off = a.rowOff ? a.rowOff[crow] : a.off;
  const float d = a.rowD ? a.rowD[crow] : a.d;
  // a row that breaks the launch's promises is left to the solve, which writes NaN for it
  if (!(L >= 1 && L <= 32 * NTL && L <= a.Lcap && M >= L && M <= a.Mcap && off >= 0 && off + M <= a.Mstride &&
        d >= 0.f))
    return;
  const int K = M - L + 1, N = TW * K;
  const int lane = threadIdx.x;
  const float2* src = a.x + ((size_t)ix * a.Nz + j) * a.Mstride + off;
  stage_rows<8, 2>(rows, MP, src, a.Mstride, TW, M, lane, 64);

  // offset of subarray n = t K + k in the staged rows, t MP + k; past the last subarray
  // (up to a whole number of steps) the zero tail of row 0 (offset M, + column < MP)
  int* const offs = reinterpret_cast<int*>(rows + TW * MP);
  const int NS = (N + 7) / 8;   // MFMA steps of 8 subarrays (16 real inner indices)
  for (int n = lane; n < 8 * NS; n += 64) {
    const int t = n / K;
    offs[n] = n < N ? t * MP + (n - t * K) : M;
  }
  __syncthreads();

  const int c = lane & 31, h = lane >> 5;
  // operands of the lane's column c + 32 q for the four subarrays of a step: BF16 high and
  // low parts of (Re, Im), and of (Im, -Re) for the imaginary part's A operand
  struct Ops {
    bf16x8 qh, ql, pih, pil;
  };
  auto operands = [&](int s8, int q) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float2 e = rows[offs[8 * s8 + 4 * h + u] + c + 32 * q];
      v[2 * u] = e.x;
      v[2 * u + 1] = e.y;
    }
    Ops o;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      o.qh[u] = (__bf16)v[u];
      o.ql[u] = (__bf16)(v[u] - (float)o.qh[u]);
    }
    // (Re, Im) -> (Im, -Re) on the BF16 bits
    const u16x8 hb = __builtin_bit_cast(u16x8, o.qh), lb = __builtin_bit_cast(u16x8, o.ql);
    u16x8 ph, pl;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ph[2 * u] = hb[2 * u + 1];
      ph[2 * u + 1] = hb[2 * u] ^ 0x8000;
      pl[2 * u] = lb[2 * u + 1];
      pl[2 * u + 1] = lb[2 * u] ^ 0x8000;
    }
    o.pih = __builtin_bit_cast(bf16x8, ph);
    o.pil = __builtin_bit_cast(bf16x8, pl);
    return o;
  };
  // block (r, q): rows of column half r, columns of half q; hi*hi + hi*lo + lo*hi
  auto block = [&](f32x16& sr, f32x16& si, const Ops& A, const Ops& B) {
    sr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.qh, B.qh, sr, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.pih, B.qh, si, 0, 0, 0);
    sr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.qh, B.ql, sr, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.pih, B.ql, si, 0, 0, 0);
    sr = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.ql, B.qh, sr, 0, 0, 0);
    si = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.pil, B.qh, si, 0, 0, 0);
  };
  f32x16 sr00 = {}, si00 = {}, sr01 = {}, si01 = {}, sr11 = {}, si11 = {};
  // the same trip count for both half-waves: every MFMA runs with the full wave
  for (int s8 = 0; s8 < NS; ++s8) {
    const Ops o0 = operands(s8, 0);
    block(sr00, si00, o0, o0);
    if constexpr (NTL == 2) {
      const Ops o1 = operands(s8, 1);
      block(sr01, si01, o0, o1);
      block(sr11, si11, o1, o1);
    }
  }

  // tr(S) = sum_t sum_m |x_t[m]|^2 #{(k, l): k + l = m, k < K, l < L}, from the data (FP32)
  float trp = 0.f;
  for (int t = 0; t < TW; ++t)
    for (int m = lane; m < M; m += 64) {
      const float2 e = rows[t * MP + m];
      const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
      trp = fmaf((float)cnt, fmaf(e.x, e.x, e.y * e.y), trp);
    }
  const float delta = d * Grp<64>::sum(trp) / (float)L;

  float2* Rp = a.R + (size_t)p * Rs;
  auto store = [&](const f32x16& sr, const f32x16& si, int r0, int c0, bool mirror) {
#pragma unroll
    for (int vv = 0; vv < 16; ++vv) {
      const int i = r0 + 8 * (vv / 4) + 4 * h + (vv % 4), cc = c0 + c;
      if (i < L && cc < L) {
        Rp[i * L + cc] = make_float2(sr[vv] + (i == cc ? delta : 0.f), -si[vv]);
        if (mirror) Rp[cc * L + i] = make_float2(sr[vv], si[vv]);
      }
    }
  };
  store(sr00, si00, 0, 0, false);
  if constexpr (NTL == 2) {
    store(sr01, si01, 0, 32, true);
    store(sr11, si11, 32, 32, false);
  }
}

hipError_t launch_cov_mfma(const CaponLaunch& a, int ntl, int Mmax, int Rs, hipStream_t s) {
  if (a.plist || !a.R || (ntl != 1 && ntl != 2) || a.Lcap > 32 * ntl || Mmax < 1) return hipErrorInvalidValue;
  const int MP = ((Mmax + 32 * ntl + 3) / 4) * 4;
  const int N = (2 * a.Y + 1) * Mmax;   // bounds every row's (2Y+1) K
  const size_t lds = (size_t)(2 * a.Y + 1) * MP * sizeof(float2) + (size_t)((N + 7) / 8 * 8) * sizeof(int);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const long blocks = a.rlist ? (long)a.npix : (long)a.Nx * a.jn;
  if (blocks == 0) return hipSuccess;
  if (ntl == 1)
    hipLaunchKernelGGL(cov_mfma_kernel<1>, dim3((unsigned)blocks), dim3(64), lds, s, a, MP, Rs);
  else
    hipLaunchKernelGGL(cov_mfma_kernel<2>, dim3((unsigned)blocks), dim3(64), lds, s, a, MP, Rs);
  return hipGetLastError();
}

}  // namespace cosk
