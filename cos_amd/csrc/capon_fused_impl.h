// Fused per-pixel Capon/MVDR kernel for gfx950: covariance build -> diagonal
// loading -> Cholesky -> forward/back solves -> power, amplitude and weights, all
// in registers of one lane group per pixel.
//
// Semantics follow the reference (paths relative to its root):
//   covariance   framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:45-161
//                S[r][c] = sum_t sum_{k<K} x_t[k+r] conj(x_t[k+c]), unnormalised,
//                loading d*tr(S)/L on the diagonal (tr = sum of real parts)
//   solve        RealTimeCapon/Solver.cpp:23-116 (here: Cholesky of S_L, a = 1)
//   outputs      RealTimeCapon/capon_kernel.cu:338-427: w = y / sum(y),
//                z = (1/K) sum_l conj(w_l) sum_k x_c[k+l]
//   power        P = 1 / (a^H R_L^-1 a) for R_L = S_L / (K(2Y+1))
//                (framework/beamformer/capon/getCaponC.pyx:240-289, intended zPow)
//
// Layout in registers.  A group of G lanes (G = 16: one DPP row, four pixels per
// wave; G = 64: one pixel per wave) owns one pixel.  Lane g owns the matrix columns
// c_s = g + G*s, s < CPL, so L <= LMAX = G*CPL.  Slot s keeps rows 0 .. G(s+1)-1 of
// its column, which covers the upper triangle (row <= column) for every lane.
// Systems with L < LMAX are padded with an identity block, so all control flow is
// uniform and per-pixel L, M and d may differ inside one wave.
//
// Covariance (FP32 path): the first row by direct sums, then the displacement
// recursion S[r][c] = S[r-1][c-1] + sum_t (x_t[K-1+r] conj(x_t[K-1+c]) -
// x_t[r-1] conj(x_t[c-1])), one DPP rotate per row, the same identity the
// reference's buildR_kernel uses along diagonals (buildR_kernel.cu:116-137).
// The FP64 path ("plain" per-pixel kernel, also the fallback for ill-conditioned
// pixels) uses direct sums for every element.
//
// Cholesky: right-looking U^H U on the upper triangle.  The multiplier U[k][i] of the
// trailing update reaches the group by a broadcast: inside the FMAs (a DPP source) for
// FP32 16-lane groups, by v_readlane for 64-lane groups, from a row published in LDS
// for most rows of the FP64 16-lane groups.  Each finished row of U is stored packed in
// LDS for the back solve, and the forward solve U^H v = a runs inside the factorisation
// (a as an extra column), so the registers of A die row by row.  a^H S_L^-1 a = |v|^2.  The back solve U y = v is
// column-oriented: one broadcast of y_j per column, U[i][j] read from LDS.
#pragma once
#include "cos_device.h"
#include "capon_kernels.h"
#include "mfma_gram.h"

namespace cosk {

template <typename T>
__device__ __forceinline__ cplx<T> ldx(const float2* p) {
  float2 v = *p;
  return {(T)v.x, (T)v.y};
}

// BWD (template argument): 0 for the forward kernel; for the adjoint, the largest M (and
// so K = M - L + 1) its epilogue handles with per-lane k and m slots (64 or 128).

// Rows of each Cholesky step whose multiplier U[k][i] is broadcast from registers
// rather than read back from LDS: every row for 64-lane groups (v_readlane into a scalar
// operand; no LDS round trip, and no registers for LDS copies), the first ten for the
// other shapes without the DPP-FMA broadcast, which covers the LDS latency of the rest.
template <int G> struct CholRegRows { static constexpr int value = G == 64 ? 64 : 10; };

// Row buffers of LMAX entries reserved after the packed U: two where row k of U is
// published in LDS for the trailing update (groups without the DPP broadcast, past the
// first CholRegRows rows) or where the 64-lane refinement passes y and beta through LDS,
// one for the w of the runtime-window kernels' options, none otherwise.
template <typename T, int G, int CPL, int TM, int BWD, bool REF> struct RowBufs {
  static constexpr bool dpp = std::is_same<T, float>::value && G == 16;
  static constexpr bool publish = !dpp && CholRegRows<G>::value < G * CPL - 1;
  static constexpr int value = (publish || (REF && G == 64)) ? 2 : (TM == 0 && !BWD) ? 1 : 0;
};

template <int G, int S, int R>
__device__ __forceinline__ constexpr int slot() { return G * S * (S + 1) / 2 + R; }

template <typename T> __device__ __forceinline__ T big_value();
template <> __device__ __forceinline__ float big_value<float>() { return 3.0e38f; }
template <> __device__ __forceinline__ double big_value<double>() { return 1.0e300; }
template <typename T> __device__ __forceinline__ T tiny_value();
template <> __device__ __forceinline__ float tiny_value<float>() { return 1.0e-30f; }
template <> __device__ __forceinline__ double tiny_value<double>() { return 1.0e-300; }

// T_: scalar type (float or double); G, CPL: lane group and columns per lane;
// TM: compile-time bound on the window length 2Y+1 (0 = runtime, reload from LDS);
// DIRECT: direct covariance sums instead of the displacement recursion.
// (1/K) sum_k x_c[k+c] exp(-j pi k sin(theta)) for one column c of a steered pixel:
// the subarray mean of the phase-rotated centre snapshot, expressed in the unrotated
// element frame (read from global memory; only the steering path uses it).
template <typename T>
__device__ __noinline__ cplx<T> steered_mean(const float2* xc, int c, int K, int L, T sn, T invK) {
  cplx<T> acc = {T(0), T(0)};
  if (c < L) {
    for (int k = 0; k < K; ++k) {
      T s2, c2;
      sincospi_t<T>(-(sn * (T)k), &s2, &c2);
      const float2 xv = xc[k + c];
      add_mul(acc, cplx<T>{(T)xv.x, (T)xv.y}, cplx<T>{(T)c2, (T)s2});
    }
  }
  return {acc.x * invK, acc.y * invK};
}

// Occupancy target (waves per SIMD) the register allocator is held to.
// (16-lane groups with one column per lane: 4, or 3 when a window of more than five
// rows keeps its alpha/beta columns in registers; three columns per lane hold 96 complex
// entries of the triangle and take the whole register file)
template <typename T, int G, int CPL, int TM> struct OccHint {
  static constexpr int value = sizeof(T) == 4 ? (G == 16 ? (CPL == 1 ? (TM == 0 || TM > 5 ? 3 : 4) : CPL == 2 ? 2 : 1) : 1)
                                              : (G == 16 && CPL == 1 ? 2 : 1);
};

// WPB: waves per block.  4 for the FP32 64-lane kernels with a fixed window bound: the
// block's waves take four pixels that, when they are consecutive rows of one beam with one
// aperture, share the staged window rows and their per-row autocorrelations (each row's
// computed once for the block instead of once per wave that needs it); a block barrier
// then hands the rows' LDS to the waves' factorisations.  1 otherwise.
// MG: for a wave whose four pixels are consecutive rows of one beam (L <= 32), 1 builds
// their covariances from per-snapshot Grams on the matrix cores (mfma_gram.h) instead of
// the displacement recursion; 2 keeps the recursion but forms its rank-two corrections on
// the matrix cores (FP16 hi/lo-split operands, FP32 accumulation, rows staged with a
// power-of-two scale) and leaves only the prefix along the diagonals to the VALU; 3 is 2
// for rows padded past 128 elements (M <= 128, padding stored as zeros), kept apart so
// that the common kernel carries no code for it.  The launch's other waves keep the VALU
// recursion.
template <typename T, int G, int CPL, int TM, bool DIRECT, int BWD = 0, bool REF = false, int MG = 0, int WPB = 1>
__device__ __forceinline__ void capon_group(const CaponLaunch& a, const int MP, const int GREG,
                                            const int gid0, const int npix) {
  using Gp = Grp<G>;
  constexpr int NG = Gp::NG * WPB, LMAX = G * CPL, NS = G * CPL * (CPL + 1) / 2;   // NG: pixels per block
  // the reference's optional variants (forward-backward averaging, power-Capon amplitude,
  // unloaded power) are compiled into the runtime-window forward kernels only, so the
  // fixed-window kernels keep their registers
  constexpr bool OPTS = TM == 0 && !BWD;
  constexpr int TMR = TM > 0 ? TM : 1;
  // one pixel per wave, FP32, fixed window bound: the recursion's rank-two sums on FP16
  // hi/lo-split matrix-core products
  constexpr bool HSPLIT = G == 64 && CPL == 1 && TM > 0 && std::is_same<T, float>::value && !BWD;
  // matrix-core recursion with rows padded past 128 elements: staged with a zero tail
  constexpr bool ZT = MG == 3 || CPL == 3;
  extern __shared__ float2 xs_all[];

  const int g = Gp::lane(), q = WPB > 1 ? (int)(threadIdx.x >> 6) : Gp::group();
  const int Y = a.Y, TW = 2 * Y + 1, Nzp = a.Nz - 2 * Y;
  const int gid = gid0 + q;
  const bool active = gid < npix;
  const int gsel = active ? gid : gid0;
  int p;
  if (a.plist) {
    p = a.plist[gsel];
  } else if (a.rlist) {
    const int nr = __builtin_amdgcn_readfirstlane(*a.rcount);   // > 0 when any group is active
    const int ixr = gsel / nr;
    p = ixr * Nzp + a.rlist[gsel - ixr * nr];
  } else {
    p = (gsel / a.jn) * Nzp + a.jlo + gsel % a.jn;
  }
  const int ix = p / Nzp, j = p - ix * Nzp, crow = j + Y;
  int L = a.rowL ? a.rowL[crow] : a.L;
  int M = a.rowM ? a.rowM[crow] : a.M;
  int off = a.rowOff ? a.rowOff[crow] : a.off;
  float dload = a.rowD ? a.rowD[crow] : a.d;
  // A row whose sizes break the launch's promises is computed as a harmless 1x1
  // problem and its outputs are written as NaN, so no access can leave the buffers.
  const bool valid = L >= 1 && L <= LMAX && L <= a.Lcap && M >= L && M <= a.Mcap && off >= 0 &&
                     off + M <= a.Mstride && dload >= 0.f;
  if (!valid) {
    L = 1;
    M = 1;
    off = 0;
    dload = 1.f;
  }
  const int K = M - L + 1;
  // 64-lane groups (one pixel per wave) run only the first L steps of the loops over
  // matrix rows; the identity padding up to LMAX is skipped instead of factorised
  constexpr bool SKIP = G == 64;
  const int Lw = SKIP ? __builtin_amdgcn_readfirstlane(L) : LMAX;

  // Shared rows: when the wave's NG pixels are consecutive in one beam with the same
  // aperture (every wave of a uniform frame whose row count is a multiple of NG), their
  // windows overlap in NG + 2Y rows, which are staged once; group q then uses rows
  // q .. q+2Y.  Otherwise each group stages its own 2Y+1 rows.
  const int NR = NG + TW - 1;
  bool shared = false, mg = false, mr = false;
  if constexpr (NG > 1 && WPB > 1) {
    // one pixel per wave: the vote goes through LDS
    __shared__ int bv[WPB][4];
    if (g == 0) {
      bv[q][0] = active ? p : -1;
      bv[q][1] = L;
      bv[q][2] = M;
      bv[q][3] = off;
    }
    __syncthreads();
    const int p0 = bv[0][0];
    bool consec = p0 >= 0 && p0 - (p0 / Nzp) * Nzp + NG <= Nzp;
    for (int w = 1; w < WPB; ++w)
      consec = consec && bv[w][0] == p0 + w && bv[w][1] == bv[0][1] && bv[w][2] == bv[0][2] && bv[w][3] == bv[0][3];
    shared = consec && (NR * MP + NR * LMAX * (int)(sizeof(cplx<T>) / sizeof(float2)) <= NG * GREG);
  } else if constexpr (NG > 1) {
    const int p0 = __builtin_amdgcn_readfirstlane(p);
    const int L0 = __builtin_amdgcn_readfirstlane(L), M0 = __builtin_amdgcn_readfirstlane(M);
    const int o0 = __builtin_amdgcn_readfirstlane(off);
    const int j0 = p0 - (p0 / Nzp) * Nzp;
    const bool mine = active && p == p0 + q && L == L0 && M == M0 && off == o0 && j0 + NG <= Nzp;
    const bool consec = __all(mine);
    shared = consec && (NR * MP + NR * LMAX * (int)(sizeof(cplx<T>) / sizeof(float2)) <= NG * GREG);
    if constexpr (MG == 1) mg = consec && L <= 32 && mfma_lds(M - L + 1, NR) <= NG * GREG;
    // (three columns per lane: the matrix cores take the block of rows and columns 0..31
    // for any L, the VALU the rest)
    if constexpr (MG >= 2)
      mr = shared && (CPL == 3 || L <= 32) && TW <= TM && NR <= 8 && (ZT ? M <= 128 : MP <= 128) &&
           NR * MP + NR * LMAX + 1024 <= NG * GREG;
  }
  float2* xs = shared ? xs_all + q * MP : xs_all + q * GREG;
  const float2* const wrow0 = a.x + ((size_t)(p - q) / Nzp * a.Nz + (j - q)) * a.Mstride + off;   // (shared, mg)
  // with mg or mr: the wave's rows were scaled by 2^pexp, so S (and P) by 2^(2 pexp)
  int pexp = 0;
  if (mg) {
    if constexpr (MG)
      pexp = mfma_stage(reinterpret_cast<unsigned*>(xs_all), wrow0, a.Mstride, __builtin_amdgcn_readfirstlane(M), NR,
                        mfma_mpf(__builtin_amdgcn_readfirstlane(M - L + 1)));
  } else if (MG >= 2 && mr) {
    pexp = stage_rows_scaled<8, 2, ZT>(xs_all, MP, wrow0, a.Mstride, NR, M);
  } else if (shared) {
    stage_rows<8, (WPB > 1 ? 1 : 2)>(xs_all, MP, wrow0, a.Mstride, NR, M, threadIdx.x, 64 * WPB);
  } else {
    // Stage the 2Y+1 snapshots of the window, zero padded to MP elements.
    stage_rows<(G == 16 ? 5 : 8), (G == 16 ? 4 : 2)>(xs, MP, a.x + ((size_t)ix * a.Nz + j) * a.Mstride + off, a.Mstride,
                                                     TW, M, g, G);
  }
  __syncthreads();

  int cs[CPL];
  static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value; cs[S] = g + G * S; });

  cplx<T> A[NS];

  // Rows R0.. of the lane's columns g and 16 + g (two columns per lane) from a window in the
  // 32 x 32 buffer (window_store's column order keeps the two side by side: one 16-byte read
  // per row for both slots)
  auto window_take = [&](auto R0_, const float2* Tb) {
    constexpr int R0 = decltype(R0_)::value;
    if constexpr (MG != 0) {
      static_assert(G == 16 && CPL >= 2, "the matrix-core covariance paths are for 16 lanes x 2 or 3 columns");
      static_for<R0, 32>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        if constexpr (R < 16) {
          const float4 v = *reinterpret_cast<const float4*>(Tb + R * 32 + 2 * g);
          A[slot<G, 0, R>()] = {(T)v.x, (T)v.y};
          A[slot<G, 1, R>()] = {(T)v.z, (T)v.w};
        } else {
          const float2 v = Tb[R * 32 + 2 * g + 1];
          A[slot<G, 1, R>()] = {(T)v.x, (T)v.y};
        }
      });
    }
  };

  if constexpr (MG == 1) {
    if (mg) {
      // pixel q's window sum of the row Grams into the transfer buffer, then into the
      // registers of group q
      // (L and M are the same in the wave's four groups)
      const unsigned* const F = reinterpret_cast<const unsigned*>(xs_all);
      const int K0 = __builtin_amdgcn_readfirstlane(M - L + 1), MPF = mfma_mpf(K0), NSTEP = (K0 + 7) >> 3;
      float2* const Tb = xs_all + 2 * NR * MPF;
      auto take = [&](const f32x16& wr, const f32x16& wi, int qq) {
        __syncthreads();   // the previous window has been read
        window_store(Tb, wr, wi);
        __syncthreads();
        if (q == qq) window_take(ic<0>{}, Tb);
      };
      const f32x16 zero = {};
      f32x16 ur, ui, xr, xi, sr, si;
      if (TW == 5) {
        // the four windows of five rows out of eight from ten Gram passes, with at most
        // two accumulator pairs live once the registers of A hold a window:
        // U = G3 + G4, X = U + G2 + G1, S0 = X + G0, S1 = X + G5,
        // Y = U + G5 + G6 (in U), S2 = Y + G2, S3 = Y + G7
        gram_pass(F, 3, MPF, NSTEP, K0, zero, zero, ur, ui);
        gram_pass(F, 4, MPF, NSTEP, K0, ur, ui, ur, ui);
        gram_pass(F, 2, MPF, NSTEP, K0, ur, ui, xr, xi);
        gram_pass(F, 1, MPF, NSTEP, K0, xr, xi, xr, xi);
        gram_pass(F, 0, MPF, NSTEP, K0, xr, xi, sr, si);
        take(sr, si, 0);
        gram_pass(F, 5, MPF, NSTEP, K0, xr, xi, xr, xi);
        take(xr, xi, 1);
        gram_pass(F, 5, MPF, NSTEP, K0, ur, ui, ur, ui);
        gram_pass(F, 6, MPF, NSTEP, K0, ur, ui, ur, ui);
        gram_pass(F, 2, MPF, NSTEP, K0, ur, ui, xr, xi);
        take(xr, xi, 2);
        gram_pass(F, 7, MPF, NSTEP, K0, ur, ui, ur, ui);
        take(ur, ui, 3);
      } else {
        for (int qq = 0; qq < NG; ++qq) {
          gram_pass(F, qq, MPF, NSTEP, K0, zero, zero, sr, si);
          for (int t = 1; t < TW; ++t) gram_pass(F, qq + t, MPF, NSTEP, K0, sr, si, sr, si);
          take(sr, si, qq);
        }
      }
    }
  }

  // ---------------- covariance S (unnormalised) ----------------
  // (or the loaded S_L of every pixel given in a.Rin, runtime-window FP32 kernels only)
  constexpr bool FR = OPTS && !REF && std::is_same<T, float>::value;
  bool from_r = false;
  if constexpr (FR) from_r = a.Rin != nullptr;
  if (mg) {
    // built above
  } else if (from_r) {
    const float2* Rp = a.Rin + (size_t)p * a.Rstride;
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      static_for<0, G * (S + 1)>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        const float2 v = (R < L && cs[S] < L) ? Rp[R * L + cs[S]] : make_float2(0.f, 0.f);
        A[slot<G, S, R>()] = {(T)v.x, (T)-v.y};
      });
    });
  } else if constexpr (DIRECT) {
    static_for<0, NS>([&](auto I_) {
    constexpr int I = decltype(I_)::value; A[I] = {T(0), T(0)}; });
    for (int t = 0; t < TW; ++t) {
      const float2* xr = xs + t * MP;
      for (int k = 0; k < K; ++k) {
        cplx<T> b[CPL];
        static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value; b[S] = ldx<T>(xr + k + cs[S]); });
        static_for<0, LMAX>([&](auto R_) {
    constexpr int R = decltype(R_)::value;
          const cplx<T> u = ldx<T>(xr + k + R);
          static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
            if constexpr (R < G * (S + 1)) add_mul_conj(A[slot<G, S, R>()], u, b[S]);
          });
        });
      }
    }
  } else {
    cplx<T> s0[CPL];
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      s0[S] = {T(0), T(0)};
    });
    if (shared) {
      // Per-row autocorrelations rho_t(c) = sum_{k<K} x_t[k] conj(x_t[k+c]) for the NR
      // shared rows, each computed once by the block (lane task: one row, NLAG lags, a
      // sliding window of NLAG values; two lags with four waves, so every lane has a task),
      // then S[0][c] = sum of the group's 2Y+1 rows.
      cplx<T>* rho = reinterpret_cast<cplx<T>*>(xs_all + NR * MP);
      constexpr int NLAG = WPB > 1 ? 2 : 4, NLB = LMAX / NLAG;
      for (int task = threadIdx.x; task < NR * NLB; task += 64 * WPB) {
        const int t = task / NLB, c0 = NLAG * (task - t * NLB);
        const float2* xr = xs_all + t * MP;
        cplx<T> acc[NLAG], wv[NLAG];
        static_for<0, NLAG>([&](auto U_) {
          constexpr int U = decltype(U_)::value;
          acc[U] = {T(0), T(0)};
          wv[U] = ldx<T>(xr + c0 + U);
        });
        // K rounded up to a multiple of NLAG; the extra terms read zero padding (a = 0)
        for (int k = 0; k < K; k += NLAG) {
          static_for<0, NLAG>([&](auto V_) {
            constexpr int V = decltype(V_)::value;
            const cplx<T> u = (k + V < K) ? ldx<T>(xr + k + V) : cplx<T>{T(0), T(0)};
            static_for<0, NLAG>([&](auto U_) {
              constexpr int U = decltype(U_)::value;
              add_mul_conj(acc[U], u, wv[(U + V) % NLAG]);
            });
            wv[V] = ldx<T>(xr + k + V + c0 + NLAG);
          });
        }
        static_for<0, NLAG>([&](auto U_) {
          constexpr int U = decltype(U_)::value;
          rho[t * LMAX + c0 + U] = acc[U];
        });
      }
      __syncthreads();
      for (int t = 0; t < TW; ++t) {
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          const cplx<T> r = rho[(q + t) * LMAX + cs[S]];
          s0[S].x += r.x;
          s0[S].y += r.y;
        });
      }
    } else {
      // eight subarrays at a time with every LDS read of the batch issued before the
      // products (read one at a time, each waited on, the loop paid the LDS latency per
      // subarray); past K the shared value is zero and the column read stays in the row
      for (int t = 0; t < TW; ++t) {
        const float2* xr = xs + t * MP;
        for (int k0 = 0; k0 < K; k0 += 8) {
          cplx<T> u[8], v[8][CPL];
          static_for<0, 8>([&](auto J_) {
            constexpr int jj = decltype(J_)::value;
            const int kk = min(k0 + jj, K - 1);
            u[jj] = ldx<T>(xr + kk);
            if (k0 + jj >= K) u[jj] = {T(0), T(0)};
            static_for<0, CPL>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              v[jj][S] = ldx<T>(xr + kk + cs[S]);
            });
          });
          static_for<0, 8>([&](auto J_) {
            constexpr int jj = decltype(J_)::value;
            static_for<0, CPL>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              add_mul_conj(s0[S], u[jj], v[jj][S]);
            });
          });
        }
      }
    }
    static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value; A[slot<G, S, 0>()] = s0[S]; });

    bool rec_done = false, hi_only = false;
    if constexpr (MG >= 2) {
      if (mr) {
        // the window sums D_q of the rank-two corrections on the matrix cores (mfma_gram.h),
        // pixel q's into rows 1.. of group q's registers through a 32 x 32 buffer, then the
        // prefix S[r][c] = S[r-1][c-1] + D[r][c] along the diagonals
        float2* const Tb = xs_all + NR * MP + NR * LMAX;
        const int K0 = __builtin_amdgcn_readfirstlane(K);
        auto take = [&](const f32x16& wr, const f32x16& wi, int qq) {
          __syncthreads();   // the previous window has been read
          window_store(Tb, wr, wi);
          __syncthreads();
          if (q == qq) window_take(ic<1>{}, Tb);
        };
        const f32x16 zero = {};
        f32x16 ur, ui, xr, xi, sr, si;
        if (TW == 5) {
          // four windows of five rows out of eight in ten passes (as the Gram tree):
          // U = E3 + E4, X = U + E2 + E1, S0 = X + E0, S1 = X + E5,
          // Y = U + E5 + E6 (in U), S2 = Y + E2, S3 = Y + E7
          rank_pass_h(xs_all, MP, 3, K0, zero, zero, ur, ui);
          rank_pass_h(xs_all, MP, 4, K0, ur, ui, ur, ui);
          rank_pass_h(xs_all, MP, 2, K0, ur, ui, xr, xi);
          rank_pass_h(xs_all, MP, 1, K0, xr, xi, xr, xi);
          rank_pass_h(xs_all, MP, 0, K0, xr, xi, sr, si);
          take(sr, si, 0);
          rank_pass_h(xs_all, MP, 5, K0, xr, xi, xr, xi);
          take(xr, xi, 1);
          rank_pass_h(xs_all, MP, 5, K0, ur, ui, ur, ui);
          rank_pass_h(xs_all, MP, 6, K0, ur, ui, ur, ui);
          rank_pass_h(xs_all, MP, 2, K0, ur, ui, xr, xi);
          take(xr, xi, 2);
          rank_pass_h(xs_all, MP, 7, K0, ur, ui, ur, ui);
          take(ur, ui, 3);
        } else {
          for (int qq = 0; qq < NG; ++qq) {
            rank_pass_h(xs_all, MP, qq, K0, zero, zero, sr, si);
            for (int t = 1; t < TW; ++t) rank_pass_h(xs_all, MP, qq + t, K0, sr, si, sr, si);
            take(sr, si, qq);
          }
        }
        static_for<1, 32>([&](auto R_) {
          constexpr int R = decltype(R_)::value;
          static_for<0, 2>([&](auto S_) {
            constexpr int S = decltype(S_)::value;
            if constexpr (R < G * (S + 1)) {
              cplx<T> own = {Gp::ror1(A[slot<G, S, R - 1>()].x), Gp::ror1(A[slot<G, S, R - 1>()].y)};
              if constexpr (S > 0 && R - 1 < G * S) {
                const cplx<T> lo = {Gp::ror1(A[slot<G, S - 1, R - 1>()].x),
                                    Gp::ror1(A[slot<G, S - 1, R - 1>()].y)};
                if (g == 0) own = lo;
              }
              A[slot<G, S, R>()].x += own.x;
              A[slot<G, S, R>()].y += own.y;
            }
          });
        });
        // three columns per lane: the third slot (columns 32..47) by the VALU recursion below
        if constexpr (CPL == 2) rec_done = true;
        else hi_only = true;
      }
    }
    if (!rec_done) {
    cplx<T> al[TMR][CPL], be[TMR][CPL];
    if constexpr (TM > 0) {
      static_for<0, TM>([&](auto t_) {
    constexpr int t = decltype(t_)::value;
        static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
          if (t < TW) {
            al[t][S] = ldx<T>(xs + t * MP + K - 1 + cs[S]);
            be[t][S] = ldx<T>(xs + t * MP + max(cs[S] - 1, 0));
          } else {
            al[t][S] = {T(0), T(0)};
            be[t][S] = {T(0), T(0)};
          }
        });
      });
    }
    if constexpr (G == 64 && CPL == 1 && TM > 0 && std::is_same<T, float>::value) {
      // One pixel per wave (FP32): the rank-two sums D = sum_t alpha_t alpha_t^H - beta_t
      // beta_t^H of all 64 rows on the matrix cores, as the three upper 32 x 32 tiles
      // (0,0), (0,1), (1,1) on v_mfma_f32_32x32x2_f32 (the operand forms of the rank-two
      // block in mfma_gram.h), then the prefix S[r][c] = S[r-1][c-1] + D[r][c] along the diagonals.
      // Lane c holds alpha_t[c] and beta_t[c]; a half-wave swap of (re, im) gives the
      // operands of both row blocks at once (rows 32P + l: re in lane l, im in lane 32 + l).
      f32x16 d00r = {}, d00i = {}, d01r = {}, d01i = {}, d11r = {}, d11i = {};
      auto swp = [](float x, float y, float& p0, float& p1) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
        p0 = __uint_as_float(r[0]);
        p1 = __uint_as_float(r[1]);
      };
      if constexpr (HSPLIT) {
        // FP16 hi/lo-split operands on v_mfma_f32_32x32x16_f16 (two of 32 cycles per tile and
        // row instead of four of 64).  The operands are scaled by 2^pexp, the largest |re|,
        // |im| of the wave's alpha and beta into [2^7, 2^8), so D is 2^(2 pexp) times the
        // data's; the first row and the trace are scaled to match below, P scaled back.
        // Inner index k = 8 (l / 32) + s: half 0 carries hi.hi and lo.lo, half 1 lo.hi and
        // hi.lo of every (re, im) pair,
        //   A (rows)    = [a_h b_h a_l b_l | a_l b_l a_h b_h]
        //   B (columns) = [a_h -b_h a_l -b_l | a_h -b_h a_l -b_l]
        // (the imaginary part's A with each pair rotated to (im, -re)).  One half-wave swap
        // of (half-0 value, half-1 value) gives the register for both row (column) blocks:
        // r[0] for rows 0..31 (lanes 32.. receive lane l - 32's value), r[1] for rows 32..63.
        float mx = 0.f;
        static_for<0, TM>([&](auto t_) {
          constexpr int t = decltype(t_)::value;
          mx = fmaxf(mx, fmaxf(fmaxf(fabsf(al[t][0].x), fabsf(al[t][0].y)), fmaxf(fabsf(be[t][0].x), fabsf(be[t][0].y))));
        });
        mx = wave_max(mx);
        int ex = 0;
        (void)frexpf(mx, &ex);   // mx < 2^ex
        pexp = (mx > 0.f && mx <= 3.0e38f) ? 8 - ex : 0;
        const float scl = ldexpf(1.f, pexp);
        auto sw2 = [](unsigned x, unsigned y) { return __builtin_amdgcn_permlane32_swap(x, y, false, false); };
        constexpr unsigned NEG = 0x80008000u;
        static_for<0, TM>([&](auto t_) {
          constexpr int t = decltype(t_)::value;
          const cplx<T> av = al[t][0];
          const cplx<T> bv = g > 0 ? be[t][0] : cplx<T>{T(0), T(0)};   // beta_t[0] = 0
          const SplitC sa = split_c(make_float2(av.x * scl, av.y * scl));
          const SplitC sb = split_c(make_float2(bv.x * scl, bv.y * scl));
          const auto x0 = sw2(sa.h, sa.l), x1 = sw2(sb.h, sb.l), x2 = sw2(sa.l, sa.h), x3 = sw2(sb.l, sb.h);
          const auto y0 = sw2(sa.hr, sa.lr), y1 = sw2(sb.hr, sb.lr), y2 = sw2(sa.lr, sa.hr), y3 = sw2(sb.lr, sb.hr);
          const unsigned nbh = sb.h ^ NEG, nbl = sb.l ^ NEG;
          const auto z0 = sw2(sa.h, sa.h), z1 = sw2(nbh, nbh), z2 = sw2(sa.l, sa.l), z3 = sw2(nbl, nbl);
          const u32x4 A0 = {x0[0], x1[0], x2[0], x3[0]}, A1 = {x0[1], x1[1], x2[1], x3[1]};
          const u32x4 I0 = {y0[0], y1[0], y2[0], y3[0]}, I1 = {y0[1], y1[1], y2[1], y3[1]};
          const u32x4 B0 = {z0[0], z1[0], z2[0], z3[0]}, B1 = {z0[1], z1[1], z2[1], z3[1]};
          d00r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(A0), as_h8(B0), d00r, 0, 0, 0);
          d01r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(A0), as_h8(B1), d01r, 0, 0, 0);
          d11r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(A1), as_h8(B1), d11r, 0, 0, 0);
          d00i = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(I0), as_h8(B0), d00i, 0, 0, 0);
          d01i = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(I0), as_h8(B1), d01i, 0, 0, 0);
          d11i = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(I1), as_h8(B1), d11i, 0, 0, 0);
        });
        A[0] = {ldexpf(A[0].x, 2 * pexp), ldexpf(A[0].y, 2 * pexp)};
      } else
      static_for<0, TM>([&](auto t_) {
        constexpr int t = decltype(t_)::value;
        const cplx<T> av = al[t][0];
        const cplx<T> bv = g > 0 ? be[t][0] : cplx<T>{T(0), T(0)};   // beta_t[0] = 0
        float ar0, ar1, ai0, ai1, br0, br1, bi0, bi1;
        swp(av.x, av.y, ar0, ar1);     // (re | im): rows 0..31, 32..63
        swp(av.y, -av.x, ai0, ai1);    // (im | -re)
        swp(bv.x, bv.y, br0, br1);
        swp(-bv.y, bv.x, bi0, bi1);    // (-im | re)
        d00r = __builtin_amdgcn_mfma_f32_32x32x2f32(ar0, ar0, d00r, 0, 0, 0);
        d01r = __builtin_amdgcn_mfma_f32_32x32x2f32(ar0, ar1, d01r, 0, 0, 0);
        d11r = __builtin_amdgcn_mfma_f32_32x32x2f32(ar1, ar1, d11r, 0, 0, 0);
        d00i = __builtin_amdgcn_mfma_f32_32x32x2f32(ai0, ar0, d00i, 0, 0, 0);
        d01i = __builtin_amdgcn_mfma_f32_32x32x2f32(ai0, ar1, d01i, 0, 0, 0);
        d11i = __builtin_amdgcn_mfma_f32_32x32x2f32(ai1, ar1, d11i, 0, 0, 0);
        d00r = __builtin_amdgcn_mfma_f32_32x32x2f32(br0, -br0, d00r, 0, 0, 0);
        d01r = __builtin_amdgcn_mfma_f32_32x32x2f32(br0, -br1, d01r, 0, 0, 0);
        d11r = __builtin_amdgcn_mfma_f32_32x32x2f32(br1, -br1, d11r, 0, 0, 0);
        d00i = __builtin_amdgcn_mfma_f32_32x32x2f32(bi0, br0, d00i, 0, 0, 0);
        d01i = __builtin_amdgcn_mfma_f32_32x32x2f32(bi0, br1, d01i, 0, 0, 0);
        d11i = __builtin_amdgcn_mfma_f32_32x32x2f32(bi1, br1, d11i, 0, 0, 0);
      });
      // accumulator element v of lane l: column 32Q + l % 32, row 32P + 8(v/4) + 4(l/32) + v%4.
      // Swapping the halves of tiles (0,0) and (0,1) gives every lane rows r0 and r0 + 4 of
      // its own column; tile (1,1) swapped with itself gives lanes 32.. rows 32 + r0 (their
      // own elements hold 32 + r0 + 4).  Lanes below 32 receive rows 32.. of columns below
      // 32, under the diagonal, which nothing reads.  Row 0 keeps the direct first row.
      static_for<0, 16>([&](auto V_) {
        constexpr int v = decltype(V_)::value;
        constexpr int r0 = 8 * (v / 4) + v % 4, r1 = r0 + 4;
        float a0, a1, b0, b1;
        swp(d00r[v], d01r[v], a0, a1);
        swp(d00i[v], d01i[v], b0, b1);
        if constexpr (r0 > 0) A[r0] = {a0, b0};
        A[r1] = {a1, b1};
        swp(d11r[v], d11r[v], a0, a1);
        swp(d11i[v], d11i[v], b0, b1);
        A[32 + r0] = {a0, b0};
        A[32 + r1] = {d11r[v], d11i[v]};
      });
      static_for<1, LMAX>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        if (R >= Lw) return;
        A[R].x += Gp::ror1(A[R - 1].x);
        A[R].y += Gp::ror1(A[R - 1].y);
      });
    } else {
    // the recursion on the VALU for the slots from S0 on (S0 = 2 after the matrix cores
    // took the first two)
    auto valu_rec = [&](auto S0_) {
    constexpr int S0 = decltype(S0_)::value;
    static_for<1, LMAX>([&](auto R_) {
    constexpr int R = decltype(R_)::value;
      if (SKIP && R >= Lw) return;
      cplx<T> nv[CPL];
      static_for<S0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
        if constexpr (R < G * (S + 1)) {
          cplx<T> own = {Gp::ror1(A[slot<G, S, R - 1>()].x), Gp::ror1(A[slot<G, S, R - 1>()].y)};
          if constexpr (S > 0 && R - 1 < G * S) {
            const cplx<T> lo = {Gp::ror1(A[slot<G, S - 1, R - 1>()].x),
                                Gp::ror1(A[slot<G, S - 1, R - 1>()].y)};
            if (g == 0) own = lo;
          }
          nv[S] = own;
        }
      });
      auto corr = [&](int t, const cplx<T>* alt, const cplx<T>* bet) {
        // one pixel per wave with the window in registers (FP32): u1 = x_t[K-1+R] and
        // u0 = x_t[R-1] are alpha_t and beta_t of lane R (read into scalars, no LDS
        // round trip; every lane's read of the row would be the same address)
        cplx<T> u1, u0;
        if constexpr (G == 64 && TM > 0 && std::is_same<T, float>::value) {
          u1 = bcastc<G, R>(alt[0]);
          u0 = bcastc<G, R>(bet[0]);
        } else {
          u1 = ldx<T>(xs + t * MP + K - 1 + R);
          u0 = ldx<T>(xs + t * MP + R - 1);
        }
        static_for<S0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
          if constexpr (R < G * (S + 1)) {
            add_mul_conj(nv[S], u1, alt[S]);
            sub_mul_conj(nv[S], u0, bet[S]);
          }
        });
      };
      if constexpr (TM > 0) {
        // branch-free over the window bound: rows past the window repeat the last row
        // against zero alpha/beta, so every LDS read of the row can be issued together
        static_for<0, TM>([&](auto t_) {
          constexpr int t = decltype(t_)::value;
          corr(t < TW ? t : TW - 1, al[t], be[t]);
        });
      } else {
        for (int t = 0; t < TW; ++t) {
          cplx<T> alt[CPL], bet[CPL];
          static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
            alt[S] = ldx<T>(xs + t * MP + K - 1 + cs[S]);
            bet[S] = ldx<T>(xs + t * MP + max(cs[S] - 1, 0));
          });
          corr(t, alt, bet);
        }
      }
      static_for<S0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
        if constexpr (R < G * (S + 1)) A[slot<G, S, R>()] = nv[S];
      });
    });
    };
    if constexpr (CPL == 3 && MG >= 2) {
      if (hi_only) valu_rec(ic<2>{});
      else valu_rec(ic<0>{});
    } else {
      valu_rec(ic<0>{});
    }
    }
    }
  }

  // ---------------- trace, loading, padding ----------------
  // tr(S) = sum_t sum_m |x_t[m]|^2 * #{(k, l): k + l = m, k < K, l < L}
  T trS;
  if (mg || mr) {
    // the diagonal of the Gram sum or of the matrix-core recursion's S (L is the same in
    // the wave's four groups; both were built from the scaled rows, as is the trace)
    T tr = T(0);
    static_for<0, LMAX>([&](auto K_) {
      constexpr int k = decltype(K_)::value;
      const T dk = Gp::template bcast<k % G>(A[slot<G, k / G, k>()].x);
      tr += k < L ? dk : T(0);   // a select: a branch per column would split the wave
    });
    trS = tr;
  } else {
    T trp = T(0);
    for (int t = 0; t < TW; ++t) {
      for (int m = g; m < M; m += G) {
        const cplx<T> v = ldx<T>(xs + t * MP + m);
        const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
        trp = fma((T)cnt, fma(v.x, v.x, v.y * v.y), trp);
      }
    }
    trS = Gp::sum(trp);
    if constexpr (HSPLIT) trS = ldexpf(trS, 2 * pexp);   // S was built 2^(2 pexp) times the data's
  }
  const T delta = from_r ? T(0) : (T)dload * trS / (T)L;   // a.Rin is loaded already
  const T invK = T(1) / (T)K;
  const bool steered = a.nsteer > 0;
  // subarray mean of the centre snapshot, g_c = (1/K) sum_k x_c[k+c], for this lane's
  // columns (computed before the rows are overwritten: early when the averaging option
  // reuses the LDS region, otherwise just before the factorisation)
  cplx<T> gS[CPL];
  auto subarray_means = [&]() {
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      cplx<T> acc = {T(0), T(0)};
      if (!steered && cs[S] < L) {
        if (MG && mg) {   // the rows are not in LDS: the centre row from memory
          const float2* xc = wrow0 + (size_t)(q + Y) * a.Mstride + cs[S];
          for (int k = 0; k < K; ++k) {
            const cplx<T> v = ldx<T>(xc + k);
            acc.x += v.x;
            acc.y += v.y;
          }
        } else {
          const float2* xc = xs + Y * MP + cs[S];
          for (int k = 0; k < K; ++k) {
            const cplx<T> v = ldx<T>(xc + k);
            acc.x += v.x;
            acc.y += v.y;
          }
        }
      }
      gS[S] = {acc.x * invK, acc.y * invK};
    });
  };
  if constexpr (OPTS) subarray_means();
  // The snapshot rows are not read past this point.
  constexpr int NU = LMAX * (LMAX + 1) / 2;

  // Forward-backward averaging S <- (S + J S^T J)/2 (CAP/getCaponAllC.pyx:357-358):
  // entry (r, c) averages with (L-1-c, L-1-r).  The upper triangle goes through the
  // group's LDS region; the trace is unchanged.
  if (OPTS && a.fb) {
    cplx<T>* const Sp = reinterpret_cast<cplx<T>*>(xs_all + q * GREG);
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      const int uc = cs[S] * (cs[S] + 1) / 2;
      static_for<0, G * (S + 1)>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        Sp[R <= cs[S] ? R + uc : NU] = A[slot<G, S, R>()];
      });
    });
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      const int rp = L - 1 - cs[S];
      static_for<0, G * (S + 1)>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        const int cp = L - 1 - R;
        const bool in = R <= cs[S] && cs[S] < L;
        const cplx<T> pv = Sp[in ? rp + cp * (cp + 1) / 2 : NU];
        cplx<T>& e = A[slot<G, S, R>()];
        if (in) e = {T(0.5) * (e.x + pv.x), T(0.5) * (e.y + pv.y)};
      });
    });
  }

  if (L < LMAX) {
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      static_for<0, G * (S + 1)>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        if (cs[S] >= L) A[slot<G, S, R>()] = {(cs[S] == R) ? T(1) : T(0), T(0)};
      });
    });
  }

  // Optional materialised R in the reference layout: mem[r*L+c] = conj(S_L[r][c]).
  if (a.R && active) {
    float2* Rp = a.R + (size_t)p * a.Rstride;
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      static_for<0, G * (S + 1)>([&](auto R_) {
        constexpr int R = decltype(R_)::value;
        const int cc = cs[S];
        if (cc < L && R <= cc) {
          cplx<T> v = A[slot<G, S, R>()];
          if (R == cc) v.x += delta;
          if constexpr (MG >= 2 || HSPLIT) v = {ldexpf(v.x, -2 * pexp), ldexpf(v.y, -2 * pexp)};
          Rp[R * L + cc] = make_float2((float)v.x, (float)-v.y);
          if (R < cc) Rp[cc * L + R] = make_float2((float)v.x, (float)v.y);
        }
      });
    });
  }

  if (a.cov_only) return;

  if constexpr (!OPTS) subarray_means();
  // every wave of the block has read the shared rows: the factorisations may overwrite them
  if constexpr (WPB > 1) __syncthreads();

  // ---------------- Cholesky U^H U = S_L, fused forward solve U^H v = a ----------------
  // LDS region of the group (free now: every read of the snapshots is complete), in
  // cplx<T> units:
  //   Upk [0, NU)           strictly upper U[r][c] (r < c) at r + c(c+1)/2; the diagonal
  //                         slot c + c(c+1)/2 holds v_c, then 0 for the back solve
  //   dump NU               target of the row-k writes of lanes whose column is <= k
  //   rowb [NU+2, NU+2+2L)  row k of U, published for the trailing update (broadcast
  //                         reads); two buffers alternate so row k+1 can be published
  //                         while row k's updates still read the other one.  Reserved
  //                         only where it is used (RowBufs): the FP32 16-lane groups
  //                         broadcast by DPP and need one buffer for the options' w
  //                         only, which keeps three columns per lane at four waves per CU
  //   invd (T) at the end   1/U_kk
  // Row k is written to Upk as it is finished, so the registers of A die row by row.
  // FP32 16-lane groups broadcast U[k][i] inside the FMAs (v_fmac_f32 with a DPP source),
  // 64-lane groups from registers; the FP64 16-lane groups publish row k in LDS for the
  // rows past the first CRR
  constexpr bool DPP_FMA = std::is_same<T, float>::value && G == 16;
  // one pixel per wave in FP32 with 64 rows: the trailing block of rows and columns 32..63
  // is updated by the first 32 steps at once on the matrix cores
  constexpr bool BLK = std::is_same<T, float>::value && G == 64 && CPL == 1;
  constexpr bool BLK2 = BLK && HSPLIT;
  // FP64, one pixel per wave: the same row blocks (the second half of each 16-row block by
  // the first half's steps, 16..31 by steps 0..15, 32..63 by 0..31, 48..63 by 32..47)
  // batched on v_mfma_f64_16x16x4f64 (dblk_update)
  constexpr bool DBLK = std::is_same<T, double>::value && G == 64 && CPL == 1;
  constexpr int CRR = CholRegRows<G>::value;
  constexpr bool PUBLISH = !DPP_FMA && CRR < LMAX - 1;
  cplx<T>* const Upk = reinterpret_cast<cplx<T>*>(xs_all + q * GREG);
  cplx<T>* const rowb = Upk + NU + 2;
  T* const invd = reinterpret_cast<T*>(xs_all + (q + 1) * GREG) - LMAX;
  int ucol[CPL];   // Upk offset of this lane's columns
  static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
    ucol[S] = cs[S] * (cs[S] + 1) / 2;
  });
  cplx<T> b1[CPL];   // right-hand side a = 1 (0 in the padding), reduced column by column
  static_for<0, CPL>([&](auto S_) {
    constexpr int S = decltype(S_)::value;
    b1[S] = {cs[S] < L ? T(1) : T(0), T(0)};
  });
  // FP32 16-lane groups with three columns per lane: the steps of row block kb (rows
  // 16kb .. 16kb+15) update only rows of that block on the VALU; the 16 x 16 blocks (ib, s)
  // of rows 16ib.. of slot s, ib > kb, take their rank-one terms on the matrix cores
  // (blk_op), in the accumulator layout from step 0 until block ib's pivots begin.  An FP32
  // MFMA occupies the SIMD for its 32 cycles (no VALU issues beside it, measured), and
  // 16 x 16 x 1 does the work of 16 FMA instructions: a win for this shape, which runs one
  // wave per SIMD and so issues at most one VALU instruction per four cycles, and neutral
  // for two columns per lane (two waves per SIMD interleave their VALU work at about 2.6
  // cycles per instruction; the headline frame took the same time either way), which keeps
  // the VALU.
  constexpr bool MBLK = DPP_FMA && CPL >= 3;
  constexpr int NBK = MBLK ? CPL * (CPL - 1) / 2 : 1;
  f32x16 bkr[NBK], bki[NBK];
  auto for_blocks = [&](auto&& f) {   // f(ic<b>, ic<ib>, ic<s>) for every block
    if constexpr (MBLK) {
      static_for<1, CPL>([&](auto IB_) {
        constexpr int ib = decltype(IB_)::value;
        static_for<ib, CPL>([&](auto S_) {
          constexpr int s = decltype(S_)::value;
          constexpr int b = (ib - 1) * CPL - (ib - 1) * ib / 2 + (s - ib);
          f(ic<b>{}, ic<ib>{}, ic<s>{});
        });
      });
    }
  };
  // BLK2: rows 16 rb + r0 .. 16 rb + 15 of every column 16 rb.. minus the rank-one terms of
  // steps m0 .. m0 + 4 nt - 1 (the second 16-row block of a half by the first's 16 steps,
  // or the second half of a 16-row block by the first half's 8), on
  // v_mfma_f32_16x16x32_f16 (one 16 x 16 column tile jb per instruction; inner index
  // k = 8 (l / 16) + s, so each 16-lane chunk g of the wave carries one step).  For four steps 4t + g at a time, a 4 x 4 transpose of 16-lane
  // chunks (xpose4) of the step rows leaves, in register jb, column tile jb's values of
  // step 4t + g in chunk g: the layout of both operands, rows (tile rb) and columns (tile
  // jb).  Each (re, im) pair is split hi + lo in FP16 with
  //   A (rows) = [h l h l],  B (columns) = [h l l h]
  // (hi.hi + lo.lo + hi.lo + lo.hi); the imaginary part's A holds the pairs rotated to
  // (im, -re), which gives -Im.  The operands are scaled by 2^eu with 2^eu sqrt(tr S_L)
  // below 2^14 (|U[m][c]|^2 <= tr S_L) and the sums applied times 2^(-2 eu).  The results
  // go back by the same transpose: register g of set r then holds, in chunk jb, row
  // 16 rb + 4g + r of column 16 jb + l % 16 (chunks below rb receive zeros; rows of the
  // tile above r0 are left as they are).
  auto sub_update = [&](auto RB_, auto R0_, auto M0_, auto NT_) {
    if constexpr (BLK2) {
      constexpr int rb = decltype(RB_)::value, r0 = decltype(R0_)::value;
      constexpr int m0 = decltype(M0_)::value, nt = decltype(NT_)::value;
      int et = 0;
      (void)frexpf((float)(trS + (T)L * delta), &et);   // tr S_L < 2^et
      const int eu = (trS > T(0) && trS < T(3.0e38)) ? 14 - ((et + 1) >> 1) : 0;
      const float su = ldexpf(1.f, eu), sre = -ldexpf(1.f, -2 * eu), sim = -sre;
      f32x4v dr[4] = {}, di[4] = {};
      static_for<0, nt>([&](auto T_) {
        constexpr int t = decltype(T_)::value;
        float xr[4], xi[4];
        static_for<0, 4>([&](auto C_) {
          constexpr int c = decltype(C_)::value;
          xr[c] = (float)A[m0 + 4 * t + c].x * su;
          xi[c] = (float)A[m0 + 4 * t + c].y * su;
        });
        xpose4(xr[0], xr[1], xr[2], xr[3]);
        xpose4(xi[0], xi[1], xi[2], xi[3]);
        const SplitC ra = split_c(make_float2(xr[rb], xi[rb]));
        const u32x4 Am = {ra.h, ra.l, ra.h, ra.l}, Im = {ra.hr, ra.lr, ra.hr, ra.lr};
        static_for<rb, 4>([&](auto J_) {
          constexpr int jb = decltype(J_)::value;
          const SplitC cb = jb == rb ? ra : split_c(make_float2(xr[jb], xi[jb]));
          const u32x4 Bm = {cb.h, cb.l, cb.l, cb.h};
          dr[jb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(Am), as_h8(Bm), dr[jb], 0, 0, 0);
          di[jb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(as_h8(Im), as_h8(Bm), di[jb], 0, 0, 0);
        });
      });
      static_for<0, 4>([&](auto R_) {
        constexpr int r = decltype(R_)::value;
        float zr[4] = {dr[0][r], dr[1][r], dr[2][r], dr[3][r]};
        float zi[4] = {di[0][r], di[1][r], di[2][r], di[3][r]};
        xpose4(zr[0], zr[1], zr[2], zr[3]);
        xpose4(zi[0], zi[1], zi[2], zi[3]);
        static_for<0, 4>([&](auto Q_) {
          constexpr int gq = decltype(Q_)::value;
          if constexpr (4 * gq + r >= r0) {   // the tile's rows above r0 are not this update's
            cplx<T>& e = A[16 * rb + 4 * gq + r];
            e.x = fma((T)zr[gq], (T)sre, e.x);
            e.y = fma((T)zi[gq], (T)sim, e.y);
          }
        });
      });
    }
  };
  // DBLK: rows 16 tr0 .. 16 tr1 - 1 (of every column from 16 tr0 on) minus the terms of
  // steps m0 .. m0 + 2 np - 1, two steps per v_mfma_f64_16x16x4f64: A[i][k] from lane
  // i + 16k, B[k][j] from lane j + 16k, D[i][j] in lane j + 16 (i % 4), element i / 4.  Per
  // step pair, X = (re, im) of rows m and m + 1 and Y = (-im, re) of the same go through a
  // 4 x 4 transpose of 16-lane chunks (xpose4d), after which register t holds, in chunk k,
  // the k-th of the four values of column 16t + l % 16: the A and B layout of tile t.  So
  // Re += X_tr . X_tc and Im += Y_tr . X_tc (the conj(U[m][i]) U[m][j] terms).  For each
  // element e of a tile row the four column tiles' results go back by the same transpose:
  // register c then holds, in chunk tc, row 16 tr + 4e + c of column 16 tc + l % 16 (zeros
  // for tiles left of the diagonal).
  auto dblk_update = [&](auto TR0_, auto TR1_, auto M0_, auto NP_, auto R0_) {
    if constexpr (DBLK) {
      // (rows of tile tr0 below 16 tr0 + r0 are not this update's)
      constexpr int tr0 = decltype(TR0_)::value, tr1 = decltype(TR1_)::value;
      constexpr int m0 = decltype(M0_)::value, np = decltype(NP_)::value, r0 = decltype(R0_)::value;
      typedef double d4 __attribute__((ext_vector_type(4)));
      d4 dre[4][4], dim[4][4];
      static_for<tr0, tr1>([&](auto R_) {
        constexpr int tr = decltype(R_)::value;
        static_for<0, 4>([&](auto C_) {
          constexpr int tc = decltype(C_)::value;
          dre[tr][tc] = d4{0.0, 0.0, 0.0, 0.0};
          dim[tr][tc] = d4{0.0, 0.0, 0.0, 0.0};
        });
      });
      static_for<0, np>([&](auto P_) {
        constexpr int m = m0 + 2 * decltype(P_)::value;
        double X[4] = {(double)A[m].x, (double)A[m].y, (double)A[m + 1].x, (double)A[m + 1].y};
        double Yv[4] = {-(double)A[m].y, (double)A[m].x, -(double)A[m + 1].y, (double)A[m + 1].x};
        xpose4d(X[0], X[1], X[2], X[3]);
        xpose4d(Yv[0], Yv[1], Yv[2], Yv[3]);
        static_for<tr0, tr1>([&](auto R_) {
          constexpr int tr = decltype(R_)::value;
          static_for<tr, 4>([&](auto C_) {
            constexpr int tc = decltype(C_)::value;
            dre[tr][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(X[tr], X[tc], dre[tr][tc], 0, 0, 0);
            dim[tr][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(Yv[tr], X[tc], dim[tr][tc], 0, 0, 0);
          });
        });
      });
      static_for<tr0, tr1>([&](auto R_) {
        constexpr int tr = decltype(R_)::value;
        static_for<0, 4>([&](auto E_) {
          constexpr int e = decltype(E_)::value;
          double qr[4], qi[4];
          static_for<0, 4>([&](auto C_) {
            constexpr int tc = decltype(C_)::value;
            qr[tc] = tc >= tr ? dre[tr][tc][e] : 0.0;
            qi[tc] = tc >= tr ? dim[tr][tc][e] : 0.0;
          });
          xpose4d(qr[0], qr[1], qr[2], qr[3]);
          xpose4d(qi[0], qi[1], qi[2], qi[3]);
          static_for<0, 4>([&](auto Q_) {
            constexpr int c = decltype(Q_)::value;
            if constexpr (tr > tr0 || 4 * e + c >= r0) {
              A[16 * tr + 4 * e + c].x -= (T)qr[c];
              A[16 * tr + 4 * e + c].y -= (T)qi[c];
            }
          });
        });
      });
    }
  };
  static_for<0, LMAX>([&](auto Kk_) {
    constexpr int k = decltype(Kk_)::value, sk = k / G, lk = k % G;
    if (SKIP && k >= Lw) return;
    if constexpr (MBLK && k == 0) {
      for_blocks([&](auto B_, auto IB_, auto S_) {
        constexpr int b = decltype(B_)::value, ib = decltype(IB_)::value, s = decltype(S_)::value;
        static_for<0, 16>([&](auto N_) {
          constexpr int n = decltype(N_)::value;
          bkr[b][n] = A[slot<G, s, 16 * ib + n>()].x;
          bki[b][n] = A[slot<G, s, 16 * ib + n>()].y;
        });
        xpose_groups(bkr[b]);
        xpose_groups(bki[b]);
      });
    }
    // A pivot that is not positive (or NaN) is clamped to tiny; the flag pass detects
    // the clamp from the largest 1/U_kk.  The loading is added in the identity padding
    // too, where it changes nothing (its right-hand side is 0).
    const T piv = fmax(Gp::template bcast<lk>(A[slot<G, sk, k>()].x) + delta, tiny_value<T>());
    const T inv = rsqrt_t<T>(piv);
    invd[k] = inv;
    cplx<T>* const rb = rowb + (k & 1) * LMAX;
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      if constexpr (k < G * (S + 1)) {
        A[slot<G, S, k>()].x *= inv;
        A[slot<G, S, k>()].y *= inv;
        if constexpr (PUBLISH) rb[cs[S]] = A[slot<G, S, k>()];
        Upk[k < cs[S] ? k + ucol[S] : NU] = A[slot<G, S, k>()];
      }
    });
    // forward step: v_k = b_k / U_kk, b_c -= conj(U[k][c]) v_k
    cplx<T> v1 = bcastc<G, lk>(b1[sk]);
    v1.x *= inv;
    v1.y *= inv;
    Upk[k + k * (k + 1) / 2] = v1;
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      if constexpr (k < G * (S + 1)) sub_conj_mul(b1[S], A[slot<G, S, k>()], v1);
    });
    // trailing update A[i][c] -= conj(U[k][i]) U[k][c]
    if constexpr (DPP_FMA) {
      // U[k][i] broadcast inside the FMAs from its owner lane.  The broadcast source, row k
      // of slot i/G, is written once at the top of the step; only the first update that
      // reads each slot needs the DPP read's wait states, the later ones follow at least one
      // block of FMAs that does not write it.
      // Row i lives in slots si .. CPL-1; they are updated two at a time sharing the
      // broadcast, an odd one alone.  With MBLK the rows past this step's row block are
      // updated on the matrix cores: four MFMAs per block, two dependent pairs.  A wave
      // issues in order, so an MFMA issued while the matrix pipe still runs its previous one
      // holds back the VALU work behind it; one is issued before each row's FMAs (at least
      // eight VALU instructions, the 32 cycles an MFMA takes), pinned there by scheduling
      // barriers, and the ones left over after the step's last row.
      constexpr int IEND = MBLK ? G * (k / G + 1) : LMAX;
      auto mfma_op = [&](auto M_) {
        constexpr int m = decltype(M_)::value;
        for_blocks([&](auto B_, auto IB_, auto S_) {
          constexpr int b = decltype(B_)::value, ib = decltype(IB_)::value, s = decltype(S_)::value;
          // blocks are numbered by ib, so the ones still live at step k come last
          if constexpr (k < 16 * ib && m / 4 == b - (NBK - nlive<CPL>(k))) {
            __builtin_amdgcn_sched_barrier(0);
            blk_op<m % 4>(bkr[b], bki[b], A[slot<G, ib, k>()].x, A[slot<G, ib, k>()].y, A[slot<G, s, k>()].x,
                          A[slot<G, s, k>()].y);
            __builtin_amdgcn_sched_barrier(0);
          }
        });
      };
      constexpr int NM = MBLK ? 4 * nlive<CPL>(k) : 0;
      static_for<k + 1, IEND>([&](auto Ii_) {
        constexpr int i = decltype(Ii_)::value, si = i / G, li = i % G;
        constexpr bool first = i == k + 1 || i % G == 0;
        constexpr int ns = CPL - si;
        if constexpr (i - k - 1 < NM) mfma_op(ic<i - k - 1>{});
        if constexpr (ns % 2 == 1)
          sub_conj_bcast_mul<li, first>(A[slot<G, si, i>()], A[slot<G, si, k>()], A[slot<G, si, k>()]);
        static_for<0, ns / 2>([&](auto P_) {
          constexpr int s0 = si + ns % 2 + 2 * decltype(P_)::value;
          sub_conj_bcast_mul2<li, first && ns % 2 == 0 && s0 == si>(A[slot<G, s0, i>()], A[slot<G, s0 + 1, i>()],
                                                                    A[slot<G, si, k>()], A[slot<G, s0, k>()],
                                                                    A[slot<G, s0 + 1, k>()]);
        });
      });
      static_for<(IEND - k - 1 < NM ? IEND - k - 1 : NM), NM>([&](auto M_) { mfma_op(M_); });
      if constexpr (MBLK) {
        // a block whose row block's pivots come next returns to the slot layout
        for_blocks([&](auto B_, auto IB_, auto S_) {
          constexpr int b = decltype(B_)::value, ib = decltype(IB_)::value, s = decltype(S_)::value;
          if constexpr (k == 16 * ib - 1) {
            xpose_groups(bkr[b]);
            xpose_groups(bki[b]);
            static_for<0, 16>([&](auto N_) {
              constexpr int n = decltype(N_)::value;
              A[slot<G, s, 16 * ib + n>()] = {bkr[b][n], bki[b][n]};
            });
          }
        });
      }
    } else {
      // the first rows take U[k][i] by a register broadcast, which covers the latency of
      // the LDS reads for the rest
      cplx<T> ur[LMAX];
      static_for<k + 1 + CRR, LMAX>([&](auto Ii_) {
        constexpr int i = decltype(Ii_)::value;
        if (SKIP && i >= Lw) return;
        ur[i] = rb[i];
      });
      auto row_update = [&](auto Ii_) {
        constexpr int i = decltype(Ii_)::value;
        const cplx<T> u = (i <= k + CRR) ? bcastc<G, i % G>(A[slot<G, i / G, k>()]) : ur[i];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          if constexpr (i < G * (S + 1)) sub_conj_mul(A[slot<G, S, i>()], u, A[slot<G, S, k>()]);
        });
      };
      // BLK: the steps of the first 32 rows update rows below 32 only; the block of rows
      // and columns 32.. takes their 32 rank-one terms at once on the matrix cores below.
      // BLK2 (the kernels on the split recursion) also leaves rows 16..31 to the first
      // 16 steps' and rows 48..63 to steps 32..47's batched terms (sub_update)
      constexpr int IEND = DBLK ? 8 * (k / 8 + 1)
                           : BLK2 ? 8 * (k / 8 + 1)
                                  : !BLK ? LMAX : k < 32 ? 32 : LMAX;
      if constexpr (SKIP) {
        // the wave steps only to its own L: rows in blocks of four with one uniform test per
        // block (a test per row cost a scalar compare and branch for every step and row);
        // rows at or past L in the last block meet U[k][i] = 0 from the identity padding
        // and change nothing
        static_for<0, (IEND - k - 1 + 3) / 4>([&](auto B_) {
          constexpr int i0 = k + 1 + 4 * decltype(B_)::value;
          if (i0 >= Lw) return;
          static_for<i0, (i0 + 4 < IEND ? i0 + 4 : IEND)>(row_update);
        });
      } else {
        static_for<k + 1, IEND>(row_update);
      }
    }
    // BLK2: the second half of each 16-row block by the first half's 8 steps, then the
    // second 16-row block of each half by the first's 16 steps (BLK takes the rest)
    if constexpr (BLK2 && k % 16 == 7) {
      if (Lw > k + 1) sub_update(ic<k / 16>{}, ic<8>{}, ic<k - 7>{}, ic<2>{});
    }
    if constexpr (BLK2 && k == 15) {
      if (Lw > 16) sub_update(ic<1>{}, ic<0>{}, ic<0>{}, ic<4>{});
    }
    if constexpr (DBLK && k % 16 == 7) {
      if (Lw > k + 1) dblk_update(ic<k / 16>{}, ic<k / 16 + 1>{}, ic<k - 7>{}, ic<4>{}, ic<8>{});
    }
    if constexpr (DBLK && k == 15) {
      if (Lw > 16) dblk_update(ic<1>{}, ic<2>{}, ic<0>{}, ic<8>{}, ic<0>{});
    }
    if constexpr (DBLK && k == 31) {
      if (Lw > 32) dblk_update(ic<2>{}, ic<4>{}, ic<0>{}, ic<16>{}, ic<0>{});
    }
    if constexpr (DBLK && k == 47) {
      if (Lw > 48) dblk_update(ic<3>{}, ic<4>{}, ic<32>{}, ic<8>{}, ic<0>{});
    }
    if constexpr (BLK2 && k == 47) {
      if (Lw > 48) sub_update(ic<3>{}, ic<0>{}, ic<32>{}, ic<4>{});
    }
    if constexpr (BLK && k == 31) {
      if (Lw > 32) {
        // S22 -= U12^H U12 for rows and columns 32..63, U12 = rows 0..31 of U in the lanes
        // of columns 32..63: per row m, Re += [Ur Ui] [Ur Ui]^T, Im += [-Ui Ur] [Ur Ui]^T on
        // v_mfma_f32_32x32x2_f32, lane l < 32 supplying column 32 + l's real part (moved
        // down by a half-wave swap) and lane 32 + l its own imaginary part.  Lanes of
        // columns below 32 receive values in rows 32.. of their columns, below the
        // diagonal, which nothing reads.
        const bool lo = (threadIdx.x & 32) == 0;
        f32x16 cr = {}, ci = {};
        float sre = 1.f, sim = 1.f;   // what the sums are multiplied by when applied
        if constexpr (HSPLIT) {
          // FP16 hi/lo-split operands, two steps per v_mfma_f32_32x32x16_f16 (32 cycles)
          // instead of one per v_mfma_f32_32x32x2_f32 (64): inner index k = 8 (l / 32) + s,
          //   A (rows)    = [h_m l_m h_m' l_m'] in both halves,
          //   B (columns) = [h_m l_m h_m' l_m'] (l < 32),  [l_m h_m l_m' h_m'] (l >= 32),
          // so one half takes hi.hi + lo.lo and the other hi.lo + lo.hi of each (re, im) pair
          // (m' = m + 1); the imaginary part's A holds the pairs rotated to (im, -re), which
          // gives -Im.  Every entry of U satisfies |U[m][c]|^2 <= S_L[c][c] <= tr S_L, so the
          // operands are scaled by 2^eu with 2^eu sqrt(tr S_L) below 2^14 (FP16 holds them
          // with room; the trace is the same in every lane); the sums are 2^(2 eu) times the
          // true ones.  Lanes of columns 32.. hold the operands; one half-wave swap hands
          // lane l < 32 those of column 32 + l.
          int et = 0;
          (void)frexpf((float)(trS + (T)L * delta), &et);   // tr S_L < 2^et
          const int eu = (trS > T(0) && trS < T(3.0e38)) ? 14 - ((et + 1) >> 1) : 0;
          const float su = ldexpf(1.f, eu);
          sre = -ldexpf(1.f, -2 * eu);
          sim = -sre;
          auto up = [](unsigned x, unsigned y) {   // lanes l < 32 get x of lane 32 + l, the rest their own y
            return __builtin_amdgcn_permlane32_swap(x, y, false, false)[1];
          };
          static_for<0, 16>([&](auto P_) {
            constexpr int m = 2 * decltype(P_)::value;
            const SplitC s0 = split_c(make_float2((float)A[m].x * su, (float)A[m].y * su));
            const SplitC s1 = split_c(make_float2((float)A[m + 1].x * su, (float)A[m + 1].y * su));
            const u32x4 Am = {up(s0.h, s0.h), up(s0.l, s0.l), up(s1.h, s1.h), up(s1.l, s1.l)};
            const u32x4 Im = {up(s0.hr, s0.hr), up(s0.lr, s0.lr), up(s1.hr, s1.hr), up(s1.lr, s1.lr)};
            const u32x4 Bm = {up(s0.h, s0.l), up(s0.l, s0.h), up(s1.h, s1.l), up(s1.l, s1.h)};
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(Am), as_h8(Bm), cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(Im), as_h8(Bm), ci, 0, 0, 0);
          });
        } else {
          sim = sre = -1.f;
          static_for<0, 32>([&](auto M_) {
            constexpr int m = decltype(M_)::value;
            const unsigned ur = __float_as_uint((float)A[m].x), ui = __float_as_uint((float)A[m].y);
            const float urd = __uint_as_float(__builtin_amdgcn_permlane32_swap(ur, ur, false, false)[1]);
            const float uid = __uint_as_float(__builtin_amdgcn_permlane32_swap(ui, ui, false, false)[1]);
            const float X = lo ? urd : (float)A[m].y;
            const float Y = lo ? -uid : (float)A[m].x;
            cr = __builtin_amdgcn_mfma_f32_32x32x2f32(X, X, cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x2f32(Y, X, ci, 0, 0, 0);
          });
        }
        // accumulator element v of lane l: column l % 32, row 8(v/4) + 4(l/32) + v%4 of the
        // block; the lower half's rows move up to the lanes of their columns
        static_for<0, 16>([&](auto V_) {
          constexpr int v = decltype(V_)::value;
          constexpr int rhi = 32 + 8 * (v / 4) + 4 + v % 4, rlo = 32 + 8 * (v / 4) + v % 4;
          const unsigned br = __float_as_uint(cr[v]), bi = __float_as_uint(ci[v]);
          const float lr = __uint_as_float(__builtin_amdgcn_permlane32_swap(br, br, false, false)[0]);
          const float li = __uint_as_float(__builtin_amdgcn_permlane32_swap(bi, bi, false, false)[0]);
          A[rhi].x = fma((T)cr[v], (T)sre, A[rhi].x);
          A[rhi].y = fma((T)ci[v], (T)sim, A[rhi].y);
          A[rlo].x = fma((T)lr, (T)sre, A[rlo].x);
          A[rlo].y = fma((T)li, (T)sim, A[rlo].y);
        });
      }
    }
  });
  // a clamped pivot gives 1/U_kk = rsqrt(tiny): the factorisation broke down
  bool notpd = false;
  if (a.flags) {
    T maxinv = T(0);
    static_for<0, LMAX>([&](auto Kk_) {
      constexpr int k = decltype(Kk_)::value;
      if (!SKIP || k < Lw) maxinv = fmax(maxinv, invd[k]);
    });
    notpd = !(maxinv < T(0.5) * rsqrt_t<T>(tiny_value<T>()));
  }

  const bool want_y = a.w || a.z || BWD || REF || (OPTS && a.pmode == 1 && a.P);
  const T norm = (T)K * (T)TW;
  const int nrhs = steered ? a.nsteer : 1;
  const float2* xcg = a.x + ((size_t)ix * a.Nz + (j + Y)) * a.Mstride + off;   // centre row (steering)

  // Forward solve U^H v = b with the rows of U read back from Upk; v_k goes to the
  // diagonal slot of column k.  Lanes whose column is <= k read the dump slot and change
  // only right-hand sides already consumed.
  int io = 0;   // 0; made opaque where the solves must not reuse earlier reads of invd
  auto fwd_upk = [&](cplx<T>(&bs)[CPL]) {
    static_for<0, LMAX>([&](auto Kk_) {
      constexpr int k = decltype(Kk_)::value, sk = k / G, lk = k % G;
      if (SKIP && k >= Lw) return;
      cplx<T> v1 = bcastc<G, lk>(bs[sk]);
      const T inv = invd[k + io];
      v1.x *= inv;
      v1.y *= inv;
      static_for<0, CPL>([&](auto S_) {
        constexpr int S = decltype(S_)::value;
        if constexpr (k < G * (S + 1)) {
          const cplx<T> u = Upk[k < cs[S] ? k + ucol[S] : NU];
          sub_conj_mul(bs[S], u, v1);
        }
      });
      Upk[k + k * (k + 1) / 2] = v1;
    });
  };
  // Back solve U y = v with v in cr, column-oriented: y_j = c_j / U_jj, then
  // c_i -= U[i][j] y_j for the rows i < j this lane owns (U[i][j] read from Upk; lanes
  // whose row is >= j read the zeroed diagonal slot of column j and leave c_i unchanged).
  auto back_upk = [&](cplx<T>(&cr)[CPL], cplx<T>(&yo)[CPL]) {
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      Upk[cs[S] + ucol[S]] = cplx<T>{T(0), T(0)};
    });
    static_for<LMAX - 1, 0, -1>([&](auto Cc_) {
      constexpr int jj = decltype(Cc_)::value, sj = jj / G, lj = jj % G;
      if (SKIP && jj >= Lw) return;
      const T inv = invd[jj + io];
      cplx<T> yb = bcastc<G, lj>(cr[sj]);
      yb.x *= inv;
      yb.y *= inv;
      static_for<0, CPL>([&](auto S_) {
        constexpr int S = decltype(S_)::value;
        if constexpr (G * S < jj) {                      // some row i = g + G*S < jj
          const cplx<T> u = Upk[min(cs[S], jj) + jj * (jj + 1) / 2];
          sub_mul(cr[S], u, yb);
        }
      });
    });
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      const T inv = cs[S] < Lw ? invd[cs[S]] : T(0);
      yo[S] = {cr[S].x * inv, cr[S].y * inv};
    });
  };

  // this block's angles: all of them, or its chunk (grid dimension y)
  const int chunk = steered && a.steer_chunk > 0 ? a.steer_chunk : nrhs;
  const int ia0 = (int)blockIdx.y * chunk, ia1 = min(nrhs, ia0 + chunk);
  for (int ia = ia0; ia < ia1; ++ia) {
    const T sn = steered ? (T)a.steer_sin[ia] : T(0);
    cplx<T> gcur[CPL];
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      gcur[S] = gS[S];
    });
    if (steered) {
      // one factorisation, many right-hand sides: forward solve for a_theta with the
      // rows of U read back from Upk
      cplx<T> bs[CPL];
      static_for<0, CPL>([&](auto S_) {
        constexpr int S = decltype(S_)::value;
        const int cc = cs[S];
        T sv, cv;
        sincospi_t<T>(sn * (T)cc, &sv, &cv);
        bs[S] = cc < L ? cplx<T>{(T)cv, (T)sv} : cplx<T>{T(0), T(0)};
        gcur[S] = steered_mean<T>(xcg, cc, K, L, sn, invK);
      });
      fwd_upk(bs);
    }

    // v (the forward solution, one entry per column) sits in the diagonal slots;
    // |v|^2 = a^H S_L^-1 a is a group sum over the lanes' columns
    cplx<T> cr[CPL];
    T accP = T(0);
    static_for<0, CPL>([&](auto S_) {
      constexpr int S = decltype(S_)::value;
      cr[S] = cs[S] < Lw ? Upk[cs[S] + ucol[S]] : cplx<T>{T(0), T(0)};
      accP = fma(cr[S].x, cr[S].x, fma(cr[S].y, cr[S].y, accP));
    });
    accP = Gp::sum(accP);

    const size_t po = (size_t)p * nrhs + ia;
    const float nanv = __builtin_nanf("");
    if (!REF && active && g == 0 && a.P && (!OPTS || a.pmode == 0)) {
      float pv = (float)(T(1) / (norm * accP));
      if constexpr (MG != 0 || HSPLIT) pv = ldexpf(pv, -2 * pexp);   // the solve ran on 2^(2 pexp) S
      a.P[po] = valid ? pv : nanv;
    }

    if (want_y) {
      cplx<T> y[CPL];
      back_upk(cr, y);
      if constexpr (REF) {
        // One step of iterative refinement against the covariance of the data itself:
        // r = a - S y - delta y, with S y = sum_t sum_k u_{t,k} beta_{t,k},
        // beta_{t,k} = u_{t,k}^H y, over the subarrays u_{t,k}[l] = x_t[k+l] of the window
        // (so the residual sees neither the recursion's nor the factorisation's rounding),
        // then S_L dy = r by the two triangular solves with U from LDS, y <- y + dy.
        // The window rows are read again from global memory (cache-resident), so the
        // refinement needs no LDS beyond the factorisation's.  Lane g takes subarray
        // k = kb + g of each block of G and loads its window xv[l] = x_t[kb + g + l],
        // l < LMAX, once: the same registers give beta_k (a dot product with y, y_l
        // broadcast inside the FMAs) and, since x_t[kb + l2 + g + G s] = xv[l2 + G s], the
        // terms x_t[k + c] beta_k of the lane's columns c = g + G s (beta_k broadcast from
        // lane l2).  Indices are clamped to the row's last active element; a clamped read
        // only meets a zero (y_l for l >= L, beta_k for k >= K) or a padding column,
        // whose residual is dropped.
        const float2* srcw = a.x + ((size_t)ix * a.Nz + j) * a.Mstride + off;
        if constexpr (G == 64) rowb[g] = y[0];   // the row buffers are dead after the factorisation
        cplx<T> sy[CPL];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          sy[S] = {T(0), T(0)};
        });
        // a short last block (K = 33 at L = 32 leaves one subarray): each beta_k as a group
        // sum over the lanes' columns, x_t[k + c] loaded once for both
        auto short_block = [&](const float2* xt, int kb) {
          for (int k = kb; k < K; ++k) {
            cplx<T> xk[CPL];
            T br = T(0), bi = T(0);
            static_for<0, CPL>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              xk[S] = ldx<T>(xt + min(k + cs[S], M - 1));
              br = fma(xk[S].x, y[S].x, fma(xk[S].y, y[S].y, br));
              bi = fma(xk[S].x, y[S].y, fma(-xk[S].y, y[S].x, bi));
            });
            const cplx<T> bk = {Gp::sum(br), Gp::sum(bi)};
            static_for<0, CPL>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              add_mul(sy[S], xk[S], bk);
            });
          }
        };
        if constexpr (G == 64) {
          for (int t = 0; t < TW; ++t) {
            const float2* xt = srcw + (size_t)t * a.Mstride;
            for (int kb = 0; kb < K; kb += G) {
              if (K - kb < G / 4) {
                short_block(xt, kb);
                break;
              }
              // one pixel per wave: y_l from LDS (one broadcast read each, nothing held in
              // scalar registers) and the window in quarters of 16 through two register
              // buffers, each quarter's loads issued one quarter ahead of its products:
              // beta over quarters 0..3, then the column terms over 3, 2 (still held), 0, 1
              // (loaded again).  Every lane takes all 64 terms with no per-term test: y_l
              // is 0 for l >= L and beta_k for k >= K (written so below), and the clamped
              // loads stay in the row, so the extra terms add exact zeros.
              // an offset the compiler cannot see through keeps the y reads inside the
              // loop (hoisted, the 64 values would hold 128 VGPRs across it)
              int yo = 0;
              asm volatile("" : "+v"(yo));
              const cplx<T>* const yl = rowb + yo;
              const float2* const xp = xt + kb + g;
              const int lim = M - 1 - kb - g;   // the row's last element, from xp
              cplx<T> xa[16], xb[16];
              cplx<T> nb = {T(0), T(0)};
              auto load_q = [&](cplx<T> (&xv)[16], auto H_) {
                constexpr int h = decltype(H_)::value;
                static_for<0, 16>([&](auto l_) {
                  constexpr int l = decltype(l_)::value;
                  xv[l] = ldx<T>(xp + min(h + l, lim));
                });
              };
              auto beta_q = [&](const cplx<T> (&xv)[16], auto H_) {
                constexpr int h = decltype(H_)::value;
                static_for<0, 16>([&](auto l_) {
                  constexpr int l = decltype(l_)::value;
                  sub_conj_mul(nb, yl[h + l], xv[l]);
                });
              };
              // beta_k reaches the column owners through LDS as well (broadcast reads)
              const cplx<T>* const bl = rowb + LMAX + yo;
              auto cols_q = [&](const cplx<T> (&xv)[16], auto H_) {
                constexpr int h = decltype(H_)::value;
                static_for<0, 16>([&](auto l_) {
                  constexpr int l2 = decltype(l_)::value;
                  sub_conj_mul(sy[0], bl[h + l2], xv[l2]);
                });
              };
              load_q(xa, ic<0>{});
              load_q(xb, ic<16>{});
              beta_q(xa, ic<0>{});
              load_q(xa, ic<32>{});
              beta_q(xb, ic<16>{});
              load_q(xb, ic<48>{});
              beta_q(xa, ic<32>{});
              beta_q(xb, ic<48>{});
              if (kb + g >= K) nb = {T(0), T(0)};
              rowb[LMAX + g] = nb;
              cols_q(xb, ic<48>{});
              load_q(xb, ic<0>{});
              cols_q(xa, ic<32>{});
              load_q(xa, ic<16>{});
              cols_q(xb, ic<0>{});
              cols_q(xa, ic<16>{});
            }
          }
        } else {
          // The full blocks of G subarrays (kb <= K - G/4) whose windows lie inside the row
          // (kb + G - 1 + LMAX <= M, so the loads take constant offsets) form one sequence
          // over all window rows, software-pipelined: the next block's window is loaded into
          // the other register buffer before the current block is computed, so its load
          // latency overlaps the FMAs (after the factorisation the triangle's registers are
          // free; one column per lane keeps a single buffer).  Past the last block the loads
          // repeat it (cache-resident, unused).  The blocks that reach the row's end follow
          // with clamped loads, then the short ones.
          const int nf = K >= G / 4 ? (K - G / 4) / G + 1 : 0;   // full blocks per row
          const int nin = min(nf, M - LMAX - G + 1 >= 0 ? (M - LMAX - G + 1) / G + 1 : 0);
          const int nblk = TW * nin;
          cplx<T> xa[LMAX], xb[LMAX];
          int lt = 0, lkb = 0;   // the next block to load
          auto load_in = [&](cplx<T> (&xv)[LMAX]) {
            const float2* const xp = srcw + (size_t)lt * a.Mstride + lkb + g;
            static_for<0, LMAX>([&](auto l_) {
              constexpr int l = decltype(l_)::value;
              xv[l] = ldx<T>(xp + l);
            });
            if (lt * nin + lkb / G + 1 < nblk) {
              lkb += G;
              if (lkb == nin * G) {
                lkb = 0;
                ++lt;
              }
            }
          };
          auto comp_blk = [&](const cplx<T> (&xv)[LMAX], int kb) {
            cplx<T> nb = {T(0), T(0)};   // -(y^H u_{t,k}) = -conj(beta_{t,k})
            // the broadcast sources (y, then nb) are written before each loop: only their
            // first reads need the DPP wait states
            static_for<0, LMAX>([&](auto l_) {
              constexpr int l = decltype(l_)::value;
              bcast_sub_conj_mul<G, l % G, l % G == 0>(nb, y[l / G], xv[l]);
            });
            if (kb + g >= K) nb = {T(0), T(0)};
            // (S y)_c += x_t[k + c] beta_{t,k} for the G subarrays of the block
            static_for<0, G>([&](auto l2_) {
              constexpr int l2 = decltype(l2_)::value;
              if constexpr (CPL >= 2)
                bcast_sub_conj_mul2<G, l2, l2 == 0>(sy[0], sy[1], nb, xv[l2], xv[l2 + G]);
              else
                bcast_sub_conj_mul<G, l2, l2 == 0>(sy[0], nb, xv[l2]);
              if constexpr (CPL == 3)   // after the pair's FMAs, which do not write nb
                bcast_sub_conj_mul<G, l2, false>(sy[2], nb, xv[l2 + 2 * G]);
            });
          };
          int ckb = 0;   // the block computed next
          auto next_ckb = [&]() {
            ckb += G;
            if (ckb == nin * G) ckb = 0;
          };
          if constexpr (CPL == 1) {   // one column per lane: 128 registers, one buffer
            for (int i = 0; i < nblk; ++i) {
              load_in(xa);
              comp_blk(xa, ckb);
              next_ckb();
            }
          } else {
            if (nblk > 0) load_in(xa);
            for (int i = 0; i < nblk; i += 2) {
              load_in(xb);
              comp_blk(xa, ckb);
              next_ckb();
              load_in(xa);
              if (i + 1 < nblk) {
                comp_blk(xb, ckb);
                next_ckb();
              }
            }
          }
          for (int t = 0; t < TW; ++t) {
            const float2* const xt = srcw + (size_t)t * a.Mstride;
            for (int kb = nin * G; kb < nf * G; kb += G) {
              static_for<0, LMAX>([&](auto l_) {
                constexpr int l = decltype(l_)::value;
                xa[l] = ldx<T>(xt + min(kb + g + l, M - 1));
              });
              comp_blk(xa, kb);
            }
            if (nf * G < K) short_block(xt, nf * G);
          }
        }
        // the right-hand side a: all ones, or the steering vector exp(j pi c sin(theta))
        cplx<T> av[CPL];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          av[S] = {T(1), T(0)};
          if (steered) {
            T sv, cv;
            sincospi_t<T>(sn * (T)cs[S], &sv, &cv);
            av[S] = {(T)cv, (T)sv};
          }
        });
        cplx<T> r[CPL];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          // matrix-core waves solved with S (and delta) 2^(2 pexp) times the data's, so y is
          // 2^(-2 pexp) times the true solution: the residual a - S_true y_true - delta_true
          // y_true takes S y from the unscaled rows times 2^(2 pexp); dy then comes out of the
          // same scaled factor, and y + dy keeps the scale
          if constexpr (MG >= 2 || HSPLIT) sy[S] = {ldexpf(sy[S].x, 2 * pexp), ldexpf(sy[S].y, 2 * pexp)};
          r[S] = cs[S] < L ? cplx<T>{av[S].x - sy[S].x - delta * y[S].x, av[S].y - sy[S].y - delta * y[S].y}
                           : cplx<T>{T(0), T(0)};
        });
        // the columns' LDS offsets and the 1/U_kk index through opaque copies: the solves
        // below recompute their addresses and read 1/U_kk again instead of holding the first
        // solves' addresses and values across the residual loop
#pragma unroll
        for (int S = 0; S < CPL; ++S) asm volatile("" : "+v"(cs[S]), "+v"(ucol[S]));
        asm volatile("" : "+v"(io));
        fwd_upk(r);
        cplx<T> cr2[CPL], dy[CPL];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          cr2[S] = cs[S] < Lw ? Upk[cs[S] + ucol[S]] : cplx<T>{T(0), T(0)};
        });
        back_upk(cr2, dy);
        T ay = T(0);
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          y[S].x += dy[S].x;
          y[S].y += dy[S].y;
          if (cs[S] < L) ay += av[S].x * y[S].x + av[S].y * y[S].y;   // Re(a^H y)
        });
        accP = Gp::sum(ay);   // Re(a^H y) of the refined solution
        if (active && g == 0 && a.P && (!OPTS || a.pmode == 0)) {
          float pv = (float)(T(1) / (norm * accP));
          if constexpr (MG != 0 || HSPLIT) pv = ldexpf(pv, -2 * pexp);
          a.P[po] = valid ? pv : nanv;
        }
      }
      // w = y / (a^H y), a^H y = |v|^2 (refined: its real part);  z = w^H g
      const T sc = T(1) / accP;
      cplx<T> zp = {T(0), T(0)};
      static_for<0, CPL>([&](auto S_) {
        constexpr int S = decltype(S_)::value;
        const int cc = cs[S];
        cplx<T> wv = {y[S].x * sc, y[S].y * sc};
        if (cc < L) add_conj_mul(zp, wv, gcur[S]);
        if (steered) {
          T sv, cv;
          sincospi_t<T>(-(sn * (T)cc), &sv, &cv);
          cplx<T> r = {T(0), T(0)};
          add_mul(r, wv, cplx<T>{(T)cv, (T)sv});
          wv = r;
        }
        if (a.w && active && cc < L) a.w[po * a.wstride + cc] = make_float2((float)wv.x, (float)wv.y);
      });
      zp.x = Gp::sum(zp.x);
      zp.y = Gp::sum(zp.y);
      if (OPTS && (a.zmode == 1 || a.pmode == 1)) {
        // w of every column, for the power-Capon amplitude and |w|^2
        cplx<T>* const wb = Upk + NU + 2;
        T wn = T(0);
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          const cplx<T> wv = cs[S] < L ? cplx<T>{y[S].x * sc, y[S].y * sc} : cplx<T>{T(0), T(0)};
          wb[cs[S]] = wv;
          wn = fma(wv.x, wv.x, fma(wv.y, wv.y, wn));
        });
        wn = Gp::sum(wn);
        if (a.pmode == 1 && active && g == 0 && a.P) {
          // |w^H R_hat w| with R_hat = S / norm unloaded: (w^H S_L w - delta |w|^2) / norm,
          // w^H S_L w = 1 / (a^H S_L^-1 a)   (CAP/getCaponMatlab_ref.m:109)
          a.P[po] = valid ? (float)(fabs(T(1) / accP - delta * wn) / norm) : nanv;
        }
        if (a.zmode == 1) {
          // power Capon: sqrt(sum_k |w^H u_{c,k}|^2) / K over the subarrays of the centre
          // snapshot (RTC/capon_kernel.cu:235-335, the compiled-out power_capon_kernel)
          T pk = T(0);
          for (int k = g; k < K; k += G) {
            cplx<T> sk = {T(0), T(0)};
            for (int l = 0; l < L; ++l) add_conj_mul(sk, wb[l], ldx<T>(xcg + k + l));
            pk = fma(sk.x, sk.x, fma(sk.y, sk.y, pk));
          }
          pk = Gp::sum(pk);
          zp = {sqrt(pk) / (T)K, T(0)};
        }
      }
      if constexpr (MG >= 2) {   // g came from the scaled rows (w is scale-free)
        if (!steered) zp = {ldexpf(zp.x, -pexp), ldexpf(zp.y, -pexp)};
      }
      if (active && g == 0 && a.z)
        a.z[po] = valid ? make_float2((float)zp.x, (float)zp.y) : make_float2(nanv, nanv);

      if constexpr (BWD) {
        // Adjoint of P = 1/(norm q), q = a^H S_L^-1 a, S_L = S + (d tr S / L) I:
        //   dq = -y^H dS y - |y|^2 d(d tr S/L),  y = S_L^-1 a
        //   dq/dx_t[m] = -2 sum_k (Gamma u_{t,k})[m-k],  Gamma = y y^H + (d |y|^2 / L) I
        //              = -2 (sum_l y_l beta_{t,m-l} + (d |y|^2/L) cnt(m) x_t[m]),
        //   beta_{t,k} = y^H u_{t,k};   dq/dd = -tr S |y|^2 / L.
        // Gradients use the convention g = d/dRe + j d/dIm (torch's for real losses).
        T yn = T(0);
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          yn = fma(y[S].x, y[S].x, fma(y[S].y, y[S].y, yn));
        });
        yn = Gp::sum(yn);
        const T gPp = (active && valid) ? (T)a.gP[p] : T(0);
        const T coeff = -gPp / (norm * accP * accP);            // dl/dq
        const T cload = (T)dload * yn / (T)L;
        // (matrix-core waves: y, accP and trS carry 2^(-2 pexp), 2^(-2 pexp) and 2^(2 pexp),
        // so coeff and the products below with y and the unscaled rows cancel theirs, and
        // only trS is taken back)
        T trS_d = trS;
        if constexpr (MG >= 2) trS_d = ldexpf(trS, -2 * pexp);
        if (active && g == 0 && a.gd) a.gd[p] = (float)(coeff * (-trS_d * yn / (T)L));
        // the window rows again (the group's region held U), then beta after them with
        // LMAX zeros in front and zeros for K <= k < KS*G, so the convolution below
        // needs no bounds tests
        constexpr int KS = (BWD + G - 1) / G;
        float2* xr0 = xs_all + q * GREG;
        cplx<T>* beta = reinterpret_cast<cplx<T>*>(xr0 + TW * MP) + LMAX;
        const float2* srcw = a.x + ((size_t)ix * a.Nz + j) * a.Mstride + off;
        for (int t = 0; t < TW; ++t)
          for (int m = g; m < MP; m += G)
            xr0[t * MP + m] = (m < M) ? srcw[(size_t)t * a.Mstride + m] : make_float2(0.f, 0.f);
        for (int i = g; i < LMAX; i += G) beta[i - LMAX] = {T(0), T(0)};
        __syncthreads();
        // lane slots k = g + G*S that hold some k < K: KA = ceil(K/G), uniform
        const int KA = (K + G - 1) / G;
        // a short last block of subarrays (K = 33 at L = 32 leaves one) is done by group
        // sums instead of a whole lane-parallel slot
        const int KF = K / G, KR = K - KF * G;
        const bool short_tail = KF >= 1 && KR > 0 && KR < G / 4;
        const int KAe = short_tail ? KF : KA;
        cplx<T> yc[CPL];
        static_for<0, CPL>([&](auto S_) {
          constexpr int S = decltype(S_)::value;
          yc[S] = {y[S].x, -y[S].y};
        });
        for (int t = 0; t < TW; ++t) {
          const float2* xt = xr0 + t * MP;
          // beta_{t,k} = y^H u_{t,k} over the KA live slots.  For k < K every read is in
          // row t (k + l <= M - 1 where y_l != 0); the lanes with k >= K read the row's
          // zero padding, the next row or the zeros in front of beta (at most G - 7
          // elements past the last row, inside the group's region), and are not kept
          static_for<1, KS + 1>([&](auto KA_) {
            constexpr int NA = decltype(KA_)::value;
            if (KAe != NA) return;
            cplx<T> nb[NA];   // -beta
            static_for<0, NA>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              nb[S] = {T(0), T(0)};
            });
            static_for<0, LMAX>([&](auto l_) {
              constexpr int l = decltype(l_)::value;
              static_for<0, (NA + 1) / 2>([&](auto P_) {
                constexpr int S = 2 * decltype(P_)::value;
                // y is written before the loop: only the first read of each slot waits
                constexpr bool nop = l % G == 0 && S == 0;
                if constexpr (S + 1 < NA)
                  bcast_sub_conj_mul2<G, l % G, nop>(nb[S], nb[S + 1], y[l / G], ldx<T>(xt + g + G * S + l),
                                                     ldx<T>(xt + g + G * (S + 1) + l));
                else
                  bcast_sub_conj_mul<G, l % G, nop>(nb[S], y[l / G], ldx<T>(xt + g + G * S + l));
              });
            });
            static_for<0, KS>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              const int k = g + G * S;
              cplx<T> v = {T(0), T(0)};
              if constexpr (S < NA) v = {-nb[S].x, -nb[S].y};
              beta[k] = k < K ? v : cplx<T>{T(0), T(0)};
            });
          });
          if (short_tail) {
            for (int k = KF * G; k < K; ++k) {   // beta_k = sum_c conj(y_c) x_t[k + c]
              T br = T(0), bi = T(0);
              static_for<0, CPL>([&](auto S_) {
                constexpr int S = decltype(S_)::value;
                const cplx<T> xv = ldx<T>(xt + k + cs[S]);
                br = fma(y[S].x, xv.x, fma(y[S].y, xv.y, br));
                bi = fma(y[S].x, xv.y, fma(-y[S].y, xv.x, bi));
              });
              br = Gp::sum(br);
              bi = Gp::sum(bi);
              if (g == k - KF * G) beta[k] = {br, bi};
            }
          }
          __syncthreads();
          // gacc_m = sum_l y_l beta_{t,m-l}, accumulated negated as -conj(conj(y_l)) beta;
          // slot S holds m < G(S+1), so l > G(S+1)-1 meets only the zero prefix and is
          // skipped at compile time
          cplx<T> ng[KS];
          static_for<0, KS>([&](auto S_) {
            constexpr int S = decltype(S_)::value;
            ng[S] = {T(0), T(0)};
          });
          static_for<0, LMAX>([&](auto l_) {
            constexpr int l = decltype(l_)::value;
            constexpr int S0 = l < G ? 0 : (l - G + 1 + G - 1) / G;   // first slot with l <= G*S + G - 1
            static_for<S0, KS, 2>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              // yc is written before the loop over t: only the first read of each slot waits
              constexpr bool nop = l % G == 0 && S == S0;
              if constexpr (S + 1 < KS)
                bcast_sub_conj_mul2<G, l % G, nop>(ng[S], ng[S + 1], yc[l / G], beta[g + G * S - l],
                                                   beta[g + G * (S + 1) - l]);
              else
                bcast_sub_conj_mul<G, l % G, nop>(ng[S], yc[l / G], beta[g + G * S - l]);
            });
          });
          cplx<T> gacc[KS];
          static_for<0, KS>([&](auto S_) {
            constexpr int S = decltype(S_)::value;
            gacc[S] = {-ng[S].x, -ng[S].y};
          });
          if (active && valid) {
            static_for<0, KS>([&](auto S_) {
              constexpr int S = decltype(S_)::value;
              const int m = g + G * S;
              if (m < M) {
                const int cnt = min(L - 1, m) - max(0, m - K + 1) + 1;
                const cplx<T> xv = ldx<T>(xt + m);
                const T f = T(-2) * coeff;
                const T vx = f * (gacc[S].x + cload * (T)cnt * xv.x);
                const T vy = f * (gacc[S].y + cload * (T)cnt * xv.y);
                float* dst = reinterpret_cast<float*>(a.gx + ((size_t)ix * a.Nz + j + t) * a.Mstride + off + m);
                atomicAdd(dst, (float)vx);
                atomicAdd(dst + 1, (float)vy);
              }
            });
          }
          __syncthreads();
        }
      }
    }
  }

  if (a.flags && active && g == 0) {
    uint8_t f = 0;
    if ((T(1) + (T)L / (T)dload) > (T)a.cond_thresh) f |= 1;
    if (notpd) f |= 2;
    if (!valid) f = 4;
    a.flags[p] = f;
  }
}

// The persistent 64-lane kernels call the per-group body out of line, so the
// grid-stride loop carries no register state across pixels (inlined, the FP64 one
// spills); the 16-lane ones keep it inline, where the call would cost registers.
template <typename T, int G, int CPL, int TM, bool DIRECT, int BWD, bool REF>
__device__ __noinline__ void capon_group_call(const CaponLaunch& a, const int MP, const int GREG,
                                              const int gid0, const int npix) {
  capon_group<T, G, CPL, TM, DIRECT, BWD, REF>(a, MP, GREG, gid0, npix);
}

// Waves per block (see capon_group): four for the FP32 64-lane forward and refinement
// kernels with a fixed window bound (the block's waves share their window rows), one
// otherwise.
template <typename T, int G, int TM, bool PERSIST, int BWD, bool REF> struct Wpb {
  static constexpr int value = (std::is_same<T, float>::value && G == 64 && TM > 0 && !PERSIST && !BWD) ? 4 : 1;
};

template <typename T, int G, int CPL, int TM, bool DIRECT, bool PERSIST, int BWD = 0, bool REF = false,
          int MG = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(64 * Wpb<T, G, TM, PERSIST, BWD, REF>::value,
                                                      64 * Wpb<T, G, TM, PERSIST, BWD, REF>::value),
                          amdgpu_waves_per_eu(BWD ? (MG ? 2 : 1)
                                              : ((REF && G == 64) || Wpb<T, G, TM, PERSIST, BWD, REF>::value > 1)
                                                  ? 2
                                                  : OccHint<T, G, CPL, TM>::value)))
void capon_kernel(const CaponLaunch a, const int MP, const int GREG) {
  // PERSIST: grid-stride over groups of NG pixels, so a launch over a device-side pixel
  // list (the FP64 fallback) can use a bounded grid; otherwise one block per group (a
  // loop costs the full-frame kernels registers).
  constexpr int WPBk = Wpb<T, G, TM, PERSIST, BWD, REF>::value;
  constexpr int NGk = Grp<G>::NG * WPBk;
  const int npix = a.pcount ? min(*a.pcount - a.pbase, a.npix) : a.rcount ? min(a.Nx * *a.rcount, a.npix) : a.npix;
  if constexpr (PERSIST) {
    for (int gid0 = blockIdx.x * NGk; gid0 < npix; gid0 += gridDim.x * NGk) {
      if constexpr (G == 64)
        capon_group_call<T, G, CPL, TM, DIRECT, BWD, REF>(a, MP, GREG, gid0, npix);
      else
        capon_group<T, G, CPL, TM, DIRECT, BWD, REF>(a, MP, GREG, gid0, npix);
    }
  } else {
    // One group of pixels per block (a loop here costs the FP32 kernel its registers).
    // Workgroups are dealt to the 8 XCDs round-robin, and each XCD has its own L2;
    // consecutive groups share window rows, so the first 8*(B/8) blocks are renumbered
    // to give each XCD a contiguous run of groups (a bijection; the tail keeps its order).
    // B counts the blocks with work: the grid of a launch over a device-side count (a row
    // list, or the fallback's pixel list) is sized for the most it may hold, and its blocks
    // past the count exit; renumbering only the live blocks keeps every XCD busy.
    constexpr unsigned NXCD = 8;
    const unsigned B = (a.rcount || a.pcount) ? (unsigned)((max(npix, 0) + NGk - 1) / NGk) : gridDim.x;
    const unsigned B8 = B / NXCD * NXCD, b = blockIdx.x;
    if (b >= B) return;
    const unsigned lb = b < B8 ? (b % NXCD) * (B8 / NXCD) + b / NXCD : b;
    const int gid0 = (int)lb * NGk;
    if (gid0 < npix) capon_group<T, G, CPL, TM, DIRECT, BWD, REF, MG, WPBk>(a, MP, GREG, gid0, npix);
  }
}


template <typename T, int G, int CPL, int TM, bool DIRECT, bool PERSIST, int BWD = 0, bool REF = false,
          int MG = 0>
hipError_t launch_one(const CaponLaunch& a, int Mmax, hipStream_t s) {
  constexpr int WPB = Wpb<T, G, TM, PERSIST, BWD, REF>::value;
  constexpr int NG = 64 / G * WPB, LMAX = G * CPL;   // pixels per block
  const int MP = ((Mmax + LMAX + 4 + 3) / 4) * 4;   // zero padding covers every lag read
  const int TW = 2 * a.Y + 1;
  // per-group LDS region: the window's snapshots, later the packed upper U + dump + pad,
  // the row buffers the variant uses and 1/U_kk (layout in capon_group)
  constexpr int NRB = RowBufs<T, G, CPL, TM, BWD, REF>::value;
  const int upk = (LMAX * (LMAX + 1) / 2 + 2 + NRB * LMAX) * (int)(sizeof(cplx<T>) / sizeof(float2)) +
                  LMAX * (int)sizeof(T) / (int)sizeof(float2);
  int GREG = TW * MP > upk ? TW * MP : upk;
  if (BWD) {   // window rows plus the zero-padded beta of the adjoint epilogue
    if (Mmax > BWD) return hipErrorInvalidValue;
    constexpr int KSG = (BWD + G - 1) / G * G;   // beta slots, after LMAX zeros
    const int need = TW * MP + (LMAX + KSG) * (int)(sizeof(cplx<T>) / sizeof(float2));
    GREG = GREG > need ? GREG : need;
  }
  const size_t lds = (size_t)NG * GREG * sizeof(float2);
  if (lds > 160 * 1024) return hipErrorInvalidValue;   // gfx950: 160 KB of LDS per workgroup
  if (!a.x) return hipErrorInvalidValue;
  int blocks = (a.npix + NG - 1) / NG;
  if (PERSIST && blocks > 2048) blocks = 2048;   // list length known only on the device
  if (blocks == 0) return hipSuccess;
  const int nchunk = a.nsteer > 0 && a.steer_chunk > 0 ? (a.nsteer + a.steer_chunk - 1) / a.steer_chunk : 1;
  hipLaunchKernelGGL((capon_kernel<T, G, CPL, TM, DIRECT, PERSIST, BWD, REF, MG>), dim3(blocks, nchunk), dim3(64 * WPB),
                     lds, s, a, MP, GREG);
  return hipGetLastError();
}

}  // namespace cosk
