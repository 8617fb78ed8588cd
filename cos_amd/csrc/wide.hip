// Capon for subarrays longer than the fused kernel's lane shapes hold (64 < L): one block of
// 256 threads per pixel, the covariance by direct FP32 sums, loading, the UᴴU Cholesky and
// both triangular solves in LDS.  It serves the drop-in entry points at the reference's own
// limits (framework/beamformer/capon/RealTimeCapon/Solver.cpp:34-42 solves up to N = 72;
// BuildR::getR builds up to L = 128), where the fused kernel stops at 64.  Semantics are
// the fused kernel's (capon_fused_impl.h): S unnormalised with loading d tr(S)/L, the R
// output conj-stored (mem[r L + c] = conj(S_L[r][c])), P = 1/(K(2Y+1) aᴴ S_L⁻¹ a) with
// a = 1, w = y / sum(y), z = wᴴ g.
//
// LDS: the window's rows [2Y+1][M] and, unless only R is asked for, S_L as a full L x L
// matrix (L <= 72: 41.5 KB), then its factor in place, the right-hand side and solution.
#include "cos_device.h"
#include "capon_kernels.h"

namespace cosk {

namespace {

constexpr int WB = 256;   // threads per pixel

__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(WB) void capon_wide_kernel(const CaponLaunch a, const int solve) {
  extern __shared__ float2 sm[];
  __shared__ float red[4];
  const int tid = threadIdx.x;
  const int Y = a.Y, TW = 2 * Y + 1, Nzp = a.Nz - 2 * Y;
  const int L = a.L, M = a.M, K = M - L + 1;
  const int p = blockIdx.x;
  const int ix = p / Nzp, j = p - ix * Nzp;
  float2* rows = sm;                          // [TW][M]
  float2* A = sm + TW * M;                    // [L][L] (row r, column c), solve only
  float2* b = A + L * L;                      // [L] right-hand side, then v, then y
  float* inv = reinterpret_cast<float*>(b + L);   // [L] 1/U_kk
  stage_rows<4, 1>(rows, M, a.x + ((size_t)ix * a.Nz + j) * a.Mstride + a.off, a.Mstride, TW, M, tid, WB);
  __syncthreads();
  // tr(S) = sum_t sum_m |x_t[m]|^2 #{(k, l): k + l = m, k < K, l < L}
  float trp = 0.f;
  for (int e = tid; e < TW * M; e += WB) {
    const int m = e % M;
    const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
    const float2 v = rows[e];
    trp = fmaf((float)cnt, fmaf(v.x, v.x, v.y * v.y), trp);
  }
  const float trS = block_sum(trp, red);
  const float delta = a.d * trS / (float)L;
  // S_L[r][c] = sum_t sum_k x_t[k+r] conj(x_t[k+c]) (+ delta on the diagonal), r <= c;
  // the R output takes both triangles, the solve the upper one
  for (int e = tid; e < L * L; e += WB) {
    const int r = e / L, c = e - r * L;
    if (r > c) continue;
    float sr = 0.f, si = 0.f;
    for (int t = 0; t < TW; ++t) {
      const float2* xr = rows + t * M;
      for (int k = 0; k < K; ++k) {
        const float2 u = xr[k + r], v = xr[k + c];
        sr = fmaf(u.x, v.x, fmaf(u.y, v.y, sr));
        si = fmaf(u.y, v.x, fmaf(-u.x, v.y, si));
      }
    }
    if (r == c) sr += delta;
    if (a.R) {
      float2* Rp = a.R + (size_t)p * a.Rstride;
      Rp[r * L + c] = make_float2(sr, -si);
      if (r < c) Rp[c * L + r] = make_float2(sr, si);
    }
    if (solve) A[r * L + c] = make_float2(sr, si);
  }
  if (!solve) return;
  for (int i = tid; i < L; i += WB) b[i] = make_float2(1.f, 0.f);
  __syncthreads();
  // right-looking Cholesky UᴴU on the upper triangle, the forward solve Uᴴv = b beside it
  for (int k = 0; k < L; ++k) {
    const float piv = fmaxf(A[k * L + k].x, 1.0e-30f);   // a non-positive pivot is clamped
    const float ik = rsqrtf(piv);
    __syncthreads();
    for (int c = k + tid; c < L; c += WB) {
      const float2 v = A[k * L + c];
      A[k * L + c] = make_float2(v.x * ik, v.y * ik);
    }
    if (tid == 0) {
      inv[k] = ik;
      b[k] = make_float2(b[k].x * ik, b[k].y * ik);   // v_k
    }
    __syncthreads();
    const float2 vk = b[k];
    const int n = L - 1 - k;   // trailing rows and columns k+1 .. L-1
    for (int e = tid; e < n * n; e += WB) {
      const int i = k + 1 + e / n, c = k + 1 + e % n;
      if (i > c) continue;
      const float2 ui = A[k * L + i], uc = A[k * L + c];
      float2& x = A[i * L + c];
      // x -= conj(ui) uc
      x.x -= ui.x * uc.x + ui.y * uc.y;
      x.y -= ui.x * uc.y - ui.y * uc.x;
    }
    for (int i = k + 1 + tid; i < L; i += WB) {           // b_i -= conj(U[k][i]) v_k
      const float2 ui = A[k * L + i];
      b[i].x -= ui.x * vk.x + ui.y * vk.y;
      b[i].y -= ui.x * vk.y - ui.y * vk.x;
    }
    __syncthreads();
  }
  // |v|^2 = aᴴ S_L⁻¹ a
  float q = 0.f;
  for (int i = tid; i < L; i += WB) q = fmaf(b[i].x, b[i].x, fmaf(b[i].y, b[i].y, q));
  q = block_sum(q, red);
  const float norm = (float)K * (float)TW;
  // back solve U y = v, column by column: y_k = v_k / U_kk, v_i -= U[i][k] y_k (i < k)
  for (int k = L - 1; k >= 0; --k) {
    __syncthreads();
    const float ik = inv[k];
    const float2 yk = make_float2(b[k].x * ik, b[k].y * ik);
    __syncthreads();
    if (tid == 0) b[k] = yk;
    for (int i = tid; i < k; i += WB) {
      const float2 u = A[i * L + k];
      b[i].x -= u.x * yk.x - u.y * yk.y;
      b[i].y -= u.x * yk.y + u.y * yk.x;
    }
  }
  __syncthreads();
  // w = y / sum(y) (sum(y) = |v|^2, real), z = wᴴ g with g_l = (1/K) sum_k x_c[k+l]
  const float sc = 1.f / q;
  float zr = 0.f, zi = 0.f;
  const float2* xc = rows + Y * M;
  for (int l = tid; l < L; l += WB) {
    const float2 w = make_float2(b[l].x * sc, b[l].y * sc);
    if (a.w) a.w[(size_t)p * a.wstride + l] = w;
    float gr = 0.f, gi = 0.f;
    for (int k = 0; k < K; ++k) {
      gr += xc[k + l].x;
      gi += xc[k + l].y;
    }
    gr /= (float)K;
    gi /= (float)K;
    zr = fmaf(w.x, gr, fmaf(w.y, gi, zr));
    zi = fmaf(w.x, gi, fmaf(-w.y, gr, zi));
  }
  zr = block_sum(zr, red);
  zi = block_sum(zi, red);
  if (tid == 0) {
    if (a.P) a.P[p] = 1.f / (norm * q);
    if (a.z) a.z[p] = make_float2(zr, zi);
  }
}

}  // namespace

hipError_t launch_capon_wide(const CaponLaunch& a, hipStream_t s) {
  if (a.L < 1 || a.L > 128 || a.M < a.L || a.rowL || a.plist || a.rlist || a.nsteer > 0) return hipErrorInvalidValue;
  const int solve = a.cov_only ? 0 : 1;
  if (solve && a.L > 72) return hipErrorInvalidValue;
  const int TW = 2 * a.Y + 1;
  const size_t lds = ((size_t)TW * a.M + (solve ? (size_t)a.L * a.L + a.L : 0)) * sizeof(float2) +
                     (solve ? (size_t)a.L * sizeof(float) : 0);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  if (a.npix <= 0) return hipSuccess;
  hipLaunchKernelGGL(capon_wide_kernel, dim3((unsigned)a.npix), dim3(WB), lds, s, a, solve);
  return hipGetLastError();
}

}  // namespace cosk
