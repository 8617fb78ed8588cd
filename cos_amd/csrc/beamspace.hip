// Beamspace Capon with a sliding DFT (sonar workloads: many beams, small Nb).
//
// Semantics (paths relative to the reference root):
//   beams      RealTimeCapon/capon_kernel.cu:47-74 -- beta(b) = b for b <= Nb/2, else
//              b + L - 2(Nb/2) - 1; B[b][l] = L^-1/2 exp(-j 2 pi beta(b) l / L)
//   transform  RealTimeCapon/sliding_beamspace_kernel.cu:26-81 -- v_{k,b} = sum_l B[b][l] x[k+l]
//              by S_{k+1} = (S_k + x[k+L] - x[k]) exp(+j 2 pi beta / L)
//   covariance R_bs = (1/(K(2Y+1))) sum_t sum_k v v^H, loading d tr(R_bs)/Nb
//              (framework/beamformer/capon/getCaponAllC.pyx:367-385, loading in beamspace)
//   solve      a_bs = B 1 = sqrt(L) e_0; w_bs = R_bs^-1 a_bs / (a_bs^H R_bs^-1 a_bs)
//   outputs    P = 1/(a_bs^H R_bs^-1 a_bs); z = w_bs^H s_c with s_c = (1/K) sum_k v_{c,k}
//              the transformed subarray mean of the centre snapshot
//              (RealTimeCapon/capon_kernel.cu:162-233); element-space w = B^H w_bs
//              (RealTimeCapon/Capon.cpp:415).
//
// Work split.  A block owns one beam and a run of RB range rows.  Phase 1: one thread per
// snapshot row runs the sliding DFT across its K subarrays and accumulates that row's
// Nb x Nb Gram and its subarray mean.  The rows are staged through LDS in windows of W
// elements (each wave loads its 64 rows with coalesced loads, then every thread walks its
// own row in the window; the next window's loads are in flight meanwhile), so the cube is
// read with few, wide requests.  Both results
// go to LDS, so each snapshot is transformed once and shared by the 2Y+1 pixels whose
// window contains it.  Phase 2: one thread per pixel sums its window's Grams, factors the
// Nb x Nb system in registers and writes the outputs.
#include "cos_device.h"
#include "beamspace.h"

namespace cosk {

template <int NB>
struct BsTri { static constexpr int N = NB * (NB + 1) / 2; };

constexpr int BS_W = 16;           // window width (elements): one 128-byte line of a row
constexpr int BS_WP = BS_W + 1;    // padded row stride in LDS (complex), spreads the banks
constexpr int BS_LD = 64 / (64 / BS_W);   // loads per lane per window (64 rows x BS_W elements)
constexpr unsigned BS_OOB = 0x80000000u;  // a buffer offset past any wave's records

// Window m0 of this wave's rows into registers: p[rr] = x[row rr*(64/BS_W) + lane/BS_W]
// [m0 + c], c = lane%BS_W, or 0 outside the row or the frame.  The loads go through a
// buffer descriptor that covers exactly the wave's valid elements, so a row past the
// frame, or a column outside [0, M) (given offset BS_OOB), reads as 0 with no branch and
// no memory access; all BS_LD loads of a window are in flight together.
__device__ __forceinline__ void bs_fetch(float2 (&p)[BS_LD], __amdgpu_buffer_rsrc_t rs, unsigned lanebase,
                                         unsigned rowstep, int M, int m0, int c) {
  const unsigned vb = (unsigned)(m0 + c) < (unsigned)M ? lanebase + (unsigned)m0 * 8u : BS_OOB;
#pragma unroll
  for (int rr = 0; rr < BS_LD; ++rr)
    p[rr] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, vb + (unsigned)rr * rowstep, 0, 0));
}

// dst[r][i] (row stride BS_WP) from the registers of bs_fetch
__device__ __forceinline__ void bs_put(float2* dst, const float2 (&p)[BS_LD], int lane) {
#pragma unroll
  for (int rr = 0; rr < BS_LD; ++rr) dst[(rr * (64 / BS_W) + lane / BS_W) * BS_WP + lane % BS_W] = p[rr];
}

// ZM: the subarray means (needed only for the amplitude z) are accumulated
template <int NB, int RB, bool ZM>
__global__ __launch_bounds__(RB) void beamspace_capon_kernel(const BeamspaceLaunch a) {
  constexpr int NT = BsTri<NB>::N;
  constexpr int NWAVE = RB / 64;
  constexpr int TILE = NWAVE * 2 * 64 * BS_WP;       // float2: tail + head window per wave
  constexpr int P2 = RB * (NT + NB);                  // float2: Grams + subarray means
  __shared__ float2 smem[TILE > P2 ? TILE : P2];
  float2 (*gram)[NT] = reinterpret_cast<float2 (*)[NT]>(smem);
  float2 (*smean)[NB] = reinterpret_cast<float2 (*)[NB]>(smem + RB * NT);
  const int Y = a.Y, TW = 2 * Y + 1, Nzp = a.Nz - 2 * Y, TP = RB - 2 * Y;
  const int ix = blockIdx.y;
  const int t0 = blockIdx.x * TP;
  const int r = threadIdx.x;
  const int L = a.L, K = a.M - a.L + 1, M = a.M;
  const float rsL = a.rsqrtL;

  // ---------------- phase 1: sliding DFT and per-snapshot Gram ----------------
  float2 G[NT];
  float2 sm[NB];
#pragma unroll
  for (int e = 0; e < NT; ++e) G[e] = make_float2(0.f, 0.f);
#pragma unroll
  for (int b = 0; b < NB; ++b) sm[b] = make_float2(0.f, 0.f);
  {
    const int wv = __builtin_amdgcn_readfirstlane(r / 64), lane = r % 64;
    const int tw0 = t0 + wv * 64;                        // first snapshot row of this wave
    const int nrows = min(64, a.Nz - tw0);               // rows of this wave inside the frame
    // descriptor over rows [tw0, tw0 + nrows), elements [0, M) of the last one; a wave
    // past the frame's end (nrows <= 0) has no records and reads zeros
    const float2* xw = a.x + ((size_t)ix * a.Nz + min(tw0, a.Nz - 1)) * a.Mstride + a.off;
    const unsigned nrec = nrows > 0 ? (unsigned)(((nrows - 1) * a.Mstride + M) * 8) : 0u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)xw, (short)0, (int)nrec, 0x00020000);
    const unsigned lanebase = (unsigned)((lane / BS_W) * a.Mstride + lane % BS_W) * 8u;
    const unsigned rowstep = (unsigned)((64 / BS_W) * a.Mstride) * 8u;
    const int col = lane % BS_W;
    float2* tail_buf = smem + wv * 2 * 64 * BS_WP;
    float2* head = tail_buf + 64 * BS_WP;
    const float2* mine_t = tail_buf + lane * BS_WP;
    const float2* mine_h = head + lane * BS_WP;
    float2 S[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) S[b] = make_float2(0.f, 0.f);
    // Windows go global -> registers one window ahead of the LDS window being walked; the
    // LDS buffers are private to the wave, whose LDS accesses complete in program order.
    // One pass over the row: S <- (S + x[l] - x[l-L]) omega for l = 0 .. M-1, x[<0] = 0.
    // From an all-zero window the first L steps build S_0[b] = sum_{l<L} x[l] exp(-j 2 pi
    // beta l / L) (omega^L = 1; the FP32 error is that of the direct sum, about 2e-6 of
    // |x|), and step l >= L-1 leaves S_k, k = l - L + 1, the sliding DFT
    // S_k = (S_{k-1} + x[k+L-1] - x[k-1]) omega.  The head windows start on multiples of
    // BS_W, so each cache line is fetched once; the tail stream re-reads the head L
    // elements behind.
    float2 om[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) om[b] = a.omega[b];
    // The L^-1/2 of B is left out of the loop: the Grams and means are scaled once, in
    // phase 2.  A diagonal Gram entry is real, so only its real part is accumulated.
    auto accumulate = [&]() {
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (!ZM) break;
        sm[b].x += S[b].x;
        sm[b].y += S[b].y;
      }
      int e = 0;
#pragma unroll
      for (int b1 = 0; b1 < NB; ++b1) {
        G[e].x = fmaf(S[b1].x, S[b1].x, fmaf(S[b1].y, S[b1].y, G[e].x));
        ++e;
#pragma unroll
        for (int b2 = b1 + 1; b2 < NB; ++b2, ++e) {   // G[b1][b2] += v[b1] conj(v[b2])
          G[e].x = fmaf(S[b1].x, S[b2].x, fmaf(S[b1].y, S[b2].y, G[e].x));
          G[e].y = fmaf(S[b1].y, S[b2].x, fmaf(-S[b1].x, S[b2].y, G[e].y));
        }
      }
    };
    auto slide = [&](float2 dlt) {
      S[0] = make_float2(S[0].x + dlt.x, S[0].y + dlt.y);   // beta(0) = 0: omega = 1
#pragma unroll
      for (int b = 1; b < NB; ++b) {
        const float2 u = make_float2(S[b].x + dlt.x, S[b].y + dlt.y);
        S[b] = make_float2(fmaf(u.x, om[b].x, -u.y * om[b].y), fmaf(u.x, om[b].y, u.y * om[b].x));
      }
    };
    float2 pt[BS_LD], ph[BS_LD];
    bs_fetch(ph, rs, lanebase, rowstep, M, 0, col);
    if (BS_W >= L) bs_fetch(pt, rs, lanebase, rowstep, M, -L, col);
    for (int p = 0; p < M; p += BS_W) {
      // the tail window [p-L, p-L+BS_W) is walked once p + BS_W >= L (zeros below 0); the
      // steps of the earlier windows are all below L - 1 and accumulate nothing
      const bool tail = p + BS_W >= L;
      bs_put(head, ph, lane);
      if (tail) bs_put(tail_buf, pt, lane);
      if (p + BS_W < M) {
        bs_fetch(ph, rs, lanebase, rowstep, M, p + BS_W, col);
        if (p + 2 * BS_W >= L) bs_fetch(pt, rs, lanebase, rowstep, M, p + BS_W - L, col);
      }
      if (tail) {
        for (int i = 0; i < BS_W && p + i < M; ++i) {
          const float2 xin = mine_h[i], xout = mine_t[i];
          slide(make_float2(xin.x - xout.x, xin.y - xout.y));
          if (p + i >= L - 1) accumulate();
        }
      } else {
        for (int i = 0; i < BS_W; ++i) slide(mine_h[i]);
      }
    }
  }
  __syncthreads();   // the windows are dead; the same LDS holds the Grams
#pragma unroll
  for (int e = 0; e < NT; ++e) gram[r][e] = G[e];
  const float invK = rsL / (float)K;                  // B's L^-1/2 and the mean's 1/K
#pragma unroll
  for (int b = 0; b < NB; ++b)
    if (ZM) smean[r][b] = make_float2(sm[b].x * invK, sm[b].y * invK);
  __syncthreads();

  // ---------------- phase 2: one pixel per thread ----------------
  const int j = t0 + r;
  if (r >= TP || j >= Nzp) return;
  float2 R[NT];
#pragma unroll
  for (int e = 0; e < NT; ++e) R[e] = make_float2(0.f, 0.f);
  for (int u = 0; u < TW; ++u) {
#pragma unroll
    for (int e = 0; e < NT; ++e) {
      const float2 gv = gram[r + u][e];
      R[e].x += gv.x;
      R[e].y += gv.y;
    }
  }
  const float nrm = 1.f / ((float)K * (float)TW * (float)L);   // with B's (L^-1/2)^2
  float tr = 0.f;
  {
    int e = 0;
#pragma unroll
    for (int b1 = 0; b1 < NB; ++b1)
#pragma unroll
      for (int b2 = b1; b2 < NB; ++b2, ++e) {
        R[e].x *= nrm;
        R[e].y *= nrm;
        if (b1 == b2) tr += R[e].x;
      }
  }
  const float dl = a.d * tr / (float)NB;
  // Cholesky U^H U = R_bs + dl I, upper triangle in place (row-major packed)
  float invd[NB];
  {
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int ii = i * NB - i * (i - 1) / 2;         // packed index of (i, i)
      float piv = R[ii].x + dl;
      piv = piv > 1e-30f ? piv : 1e-30f;
      const float inv = __builtin_amdgcn_rsqf(piv);
      invd[i] = inv;
#pragma unroll
      for (int c = i + 1; c < NB; ++c) {
        R[ii + c - i].x *= inv;
        R[ii + c - i].y *= inv;
      }
#pragma unroll
      for (int i2 = i + 1; i2 < NB; ++i2) {
        const float2 ui = R[ii + i2 - i];                 // U[i][i2]
        const int rr = i2 * NB - i2 * (i2 - 1) / 2;       // packed (i2, i2)
#pragma unroll
        for (int c = i2; c < NB; ++c) {                   // A[i2][c] -= conj(U[i][i2]) U[i][c]
          const float2 uc = R[ii + c - i];
          R[rr + c - i2].x -= ui.x * uc.x + ui.y * uc.y;
          R[rr + c - i2].y -= ui.x * uc.y - ui.y * uc.x;
        }
      }
    }
  }
  // forward U^H v = sqrt(L) e_0
  float2 v[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    float2 acc = make_float2(i == 0 ? a.sqrtL : 0.f, 0.f);
#pragma unroll
    for (int k = 0; k < i; ++k) {                       // acc -= conj(U[k][i]) v_k
      const float2 u = R[k * NB - k * (k - 1) / 2 + i - k];
      acc.x -= u.x * v[k].x + u.y * v[k].y;
      acc.y -= u.x * v[k].y - u.y * v[k].x;
    }
    v[i] = make_float2(acc.x * invd[i], acc.y * invd[i]);
  }
  float q = 0.f;
#pragma unroll
  for (int i = 0; i < NB; ++i) q += v[i].x * v[i].x + v[i].y * v[i].y;
  const size_t p = (size_t)ix * Nzp + j;
  if (a.P) a.P[p] = 1.f / q;
  if (!(a.z || a.wbs || a.w)) return;
  // back U y = v
  float2 y[NB];
#pragma unroll
  for (int i = NB - 1; i >= 0; --i) {
    float2 acc = v[i];
    const int ii = i * NB - i * (i - 1) / 2;
#pragma unroll
    for (int c = i + 1; c < NB; ++c) {                  // acc -= U[i][c] y_c
      const float2 u = R[ii + c - i];
      acc.x -= u.x * y[c].x - u.y * y[c].y;
      acc.y -= u.x * y[c].y + u.y * y[c].x;
    }
    y[i] = make_float2(acc.x * invd[i], acc.y * invd[i]);
  }
  const float iq = 1.f / q;
  float2 zz = make_float2(0.f, 0.f);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const float2 wb = make_float2(y[b].x * iq, y[b].y * iq);
    if (ZM) {
      const float2 s = smean[r + Y][b];
      zz.x += wb.x * s.x + wb.y * s.y;                  // conj(w) s
      zz.y += wb.x * s.y - wb.y * s.x;
    }
    y[b] = wb;
    if (a.wbs) a.wbs[p * NB + b] = wb;
  }
  if (ZM && a.z) a.z[p] = zz;
  if (a.w) {
    for (int l = 0; l < L; ++l) {                       // w = B^H w_bs
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const float2 tw = a.tw[b * L + l];              // B[b][l] = tw / sqrt(L)
        acc.x += tw.x * y[b].x + tw.y * y[b].y;
        acc.y += tw.x * y[b].y - tw.y * y[b].x;
      }
      a.w[p * a.wstride + l] = make_float2(acc.x * rsL, acc.y * rsL);
    }
  }
}

template <int NB>
static hipError_t launch_bs(const BeamspaceLaunch& a, hipStream_t s) {
  constexpr int RB = 128;
  const int TP = RB - 2 * a.Y;
  if (TP < 1) return hipErrorInvalidValue;
  const int Nzp = a.Nz - 2 * a.Y;
  dim3 grid((Nzp + TP - 1) / TP, a.Nx);
  if (a.z)
    hipLaunchKernelGGL((beamspace_capon_kernel<NB, RB, true>), grid, dim3(RB), 0, s, a);
  else
    hipLaunchKernelGGL((beamspace_capon_kernel<NB, RB, false>), grid, dim3(RB), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_beamspace(const BeamspaceLaunch& a, hipStream_t s) {
  switch (a.Nb) {
    case 1: return launch_bs<1>(a, s);
    case 2: return launch_bs<2>(a, s);
    case 3: return launch_bs<3>(a, s);
    case 4: return launch_bs<4>(a, s);
    case 5: return launch_bs<5>(a, s);
    case 6: return launch_bs<6>(a, s);
    case 7: return launch_bs<7>(a, s);
    case 8: return launch_bs<8>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cosk
