// Per-snapshot Gram matrices on the matrix cores, for the fused Capon kernel.
//
// A wave of the fused kernel holds four pixels that are consecutive range rows of one beam
// (16 lanes each).  Their windows cover NR = 4 + 2Y snapshot rows, and pixel q's covariance
// is the sum of the Grams of rows q .. q+2Y:
//   S_q = sum_{t=q}^{q+2Y} G_t,   G_t[i][j] = sum_{k<K} x_t[k+i] conj(x_t[k+j])
// (framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:45-161 forms the same S per
// pixel).  Each G_t is a 32 x 32 complex product over the row's K subarrays, computed once
// per wave on v_mfma_f32_32x32x16_f16 and shared by every window that contains the row,
// instead of the displacement recursion on the VALU.
//
// Precision.  One power-of-two scale per wave, 2^s with s chosen so that the largest
// |Re|, |Im| of the wave's rows lies in [2^7, 2^8), keeps every value inside the FP16
// range; each scaled value is split x = hi + lo with hi = fp16(x), lo = fp16(x - hi)
// (round to nearest even, 22 significant bits together), and a Gram is hi*hi + hi*lo +
// lo*hi accumulated in FP32 (the omitted lo*lo is 2^-22 of the product).  The scale is
// exact: the solve runs on 2^(2s) S, whose weights and amplitude equal those of S, and its
// power is scaled back by 2^(-2s).
//
// Operand layout (v_mfma_f32_32x32x16_f16, wave64): lane l supplies A[i = l%32][k] and
// B[k][j = l%32] for k = 8(l/32) + 0..7, and holds D[i][j] for j = l%32,
// i = 8(v/4) + 4(l/32) + v%4 in accumulator element v < 16.  With the inner index
// k = 2u + part over four subarrays n = 8s + 4(l/32) + u of step s, the real part of the
// Gram is Q^T Q and the imaginary part P^T Q, Q[k][c] = (Re, Im)(x[n + c]) and
// P[k][c] = (Im, -Re)(x[n + c]): each lane reads the four consecutive elements
// x[8s + 4(l/32) + c .. +3] of its column c, for A and B alike.  The split rows are staged
// in LDS once per wave as four planes per row, Qhi, Qlo, Phi, Plo (one 32-bit pair of
// halves per element), so each operand is one contiguous 16-byte read into consecutive
// registers: a step costs four LDS reads and six MFMAs, and no VALU work.
#pragma once
#include "cos_device.h"

namespace cosk {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Split-row length per snapshot row (elements): every read x[n + c] with n < 8 ceil(K/8),
// c < 32 lies below it.
__device__ __forceinline__ int mfma_mpf(int K) { return 8 * ((K + 7) >> 3) + 32; }
// LDS the path needs for a wave, in float2 units: the split rows [NR][4 planes][MPF] of
// 4 bytes and the 32 x 32 window buffer.
__device__ __forceinline__ int mfma_lds(int K, int NR) { return NR * mfma_mpf(K) * 2 + 1024; }

__device__ __forceinline__ unsigned pack_h2(_Float16 lo, _Float16 hi) {
  return (unsigned)__builtin_bit_cast(unsigned short, lo) | ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
}

// Qhi, Qlo, Phi, Plo of one scaled element (re, im) into the four planes of its row
__device__ __forceinline__ void split_f16(unsigned* row, int MPF, int m, float re, float im) {
  const _Float16 rh = (_Float16)re, ih = (_Float16)im;
  const _Float16 rl = (_Float16)(re - (float)rh), il = (_Float16)(im - (float)ih);
  row[m] = pack_h2(rh, ih);
  row[MPF + m] = pack_h2(rl, il);
  row[2 * MPF + m] = pack_h2(ih, -rh);
  row[3 * MPF + m] = pack_h2(il, -rl);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Stages the wave's NR rows (src0 + t Mstride, M elements each) as split FP16 planes
// F[t][plane][m], m < MPF, zero past M.  Returns the exponent e of the power scale applied
// to the rows: the data's S is 2^(-2e) times the Gram sums built from F.
__device__ __forceinline__ int mfma_stage(unsigned* F, const float2* src0, int Mstride, int M, int NR, int MPF) {
  float mx = 0.f;
  for (int t = 0; t < NR; ++t)
    for (int m = threadIdx.x; m < M; m += 64) {
      const float2 v = src0[(size_t)t * Mstride + m];
      mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
  mx = wave_max(mx);
  int ex = 0;
  (void)frexpf(mx, &ex);                       // mx < 2^ex
  const int e = (mx > 0.f && mx <= 3.0e38f) ? 8 - ex : 0;
  const float sc = ldexpf(1.f, e);
  for (int t = 0; t < NR; ++t)
    for (int m = threadIdx.x; m < MPF; m += 64) {
      float2 v = make_float2(0.f, 0.f);
      if (m < M) v = src0[(size_t)t * Mstride + m];
      split_f16(F + 4 * t * MPF, MPF, m, v.x * sc, v.y * sc);
    }
  return e;
}

__device__ __forceinline__ half8 as_h8(u32x4 v) { return __builtin_bit_cast(half8, v); }

__device__ __forceinline__ u32x4 ld4(const unsigned* p) {
  u32x4 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}

// One step: the lane's four subarrays n = 8 s8 + 4h + u into (r, im); in the last step
// (MASK) the B operands of subarrays n >= K are zero.
template <bool MASK>
__device__ __forceinline__ void gram_step(const unsigned* Ft, int MPF, int s8, int K, const f32x16& cr,
                                          const f32x16& ci, f32x16& r, f32x16& im) {
  const unsigned* e = Ft + 8 * s8;
  u32x4 qh = ld4(e), ql = ld4(e + MPF);
  const u32x4 ph = ld4(e + 2 * MPF), pl = ld4(e + 3 * MPF);
  if constexpr (MASK) {
    const int nv = K - (8 * s8 + 4 * ((threadIdx.x >> 5) & 1));
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      qh[u] = u < nv ? qh[u] : 0u;
      ql[u] = u < nv ? ql[u] : 0u;
    }
  }
  r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(qh), as_h8(qh), cr, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(ph), as_h8(qh), ci, 0, 0, 0);
  r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(qh), as_h8(ql), r, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(ph), as_h8(ql), im, 0, 0, 0);
  r = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(ql), as_h8(qh), r, 0, 0, 0);
  im = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(pl), as_h8(qh), im, 0, 0, 0);
}

// D = C + G_t: the Gram of staged row t (K subarrays, NSTEP = ceil(K/8) steps, both
// wave-uniform), real part in dr, imaginary in di.  The first step reads C, so a chain
// that starts from another accumulator copies nothing.
__device__ __forceinline__ void gram_pass(const unsigned* F, int t, int MPF, int NSTEP, int K, const f32x16& cr,
                                          const f32x16& ci, f32x16& dr, f32x16& di) {
  const int c = threadIdx.x & 31, h = (threadIdx.x >> 5) & 1;
  const unsigned* Ft = F + 4 * t * MPF + 4 * h + c;
  const bool ragged = 8 * NSTEP > K;
  f32x16 r, im;
  if (NSTEP == 1) {
    if (ragged) gram_step<true>(Ft, MPF, 0, K, cr, ci, r, im);
    else gram_step<false>(Ft, MPF, 0, K, cr, ci, r, im);
  } else {
    gram_step<false>(Ft, MPF, 0, K, cr, ci, r, im);
    for (int s8 = 1; s8 < NSTEP - 1; ++s8) gram_step<false>(Ft, MPF, s8, K, r, im, r, im);
    if (ragged) gram_step<true>(Ft, MPF, NSTEP - 1, K, r, im, r, im);
    else gram_step<false>(Ft, MPF, NSTEP - 1, K, r, im, r, im);
  }
  dr = r;
  di = im;
}

// ---- rank-two corrections of the displacement recursion (matrix cores) ---------------
// The recursion S[r][c] = S[r-1][c-1] + D[r][c] (capon_fused_impl.h) adds, per window row t,
//   E_t[r][c] = alpha_t[r] conj(alpha_t[c]) - beta_t[r] conj(beta_t[c]),
//   alpha_t[r] = x_t[K-1+r], beta_t[r] = x_t[r-1]
// (framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:116-137 forms the same terms
// element by element).  For a 32 x 32 block, in real terms,
//   Re E = [ar ai] [ar ai]^T - [br bi] [br bi]^T
//   Im E = [ai -ar] [ar ai]^T + [-bi br] [br bi]^T
// with lane l supplying row (and column) l % 32.  A window sum of such terms replaces the
// recursion's VALU work; the prefix along the diagonals stays on the VALU.  With FP32
// operands each product is one v_mfma_f32_32x32x2_f32 (inner index l / 32), 64 cycles
// during which the SIMD issues nothing else; rank_pass_h puts all four, the operands split
// hi + lo in FP16, into the inner index of two v_mfma_f32_32x32x16_f16 of 32 cycles each
// (headline frame: 1.383 -> 1.329 ms, P error against FP64 max 5.1e-6 -> 4.8e-6).  (Three
// 16 x 16 blocks on v_mfma_f32_16x16x4_f32, six independent chains per row instead of two,
// measured slower: 1.617 against 1.536 ms per frame, the operand selection for four inner
// indices costing more VALU than the shorter chains saved.)

// Stages the wave's nrows <= RB rows of M <= 64 CB elements (src + t sstride, zero padded
// to MP) scaled by 2^e, e chosen so that the largest |re| or |im| lies in [2^7, 2^8)
// (e = 0 for all-zero or non-finite data), and returns e: a power-of-two scale is exact,
// so everything built from the rows is 2^e (or 2^(2e) for second moments) times the
// unscaled result, and the FP16 split below keeps its 22 significant bits for every
// value down to 2^-14 of the largest.  All loads are issued before the reduction; with
// ZTAIL, padding past 64 CB is stored as zeros without loads (otherwise MP <= 64 CB).
template <int RB, int CB, bool ZTAIL = false>
__device__ __forceinline__ int stage_rows_scaled(float2* dst, int MP, const float2* src, size_t sstride, int nrows,
                                                 int M) {
  const int l = threadIdx.x & 63;
  float2 v[RB][CB];
  float mx = 0.f;
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int m = l + 64 * c;
      // every lane loads (a clamped, valid address) and selects: no branch per lane
      const float2 t = src[(size_t)min(r, nrows - 1) * sstride + min(m, M - 1)];
      v[r][c] = (r < nrows && m < M) ? t : make_float2(0.f, 0.f);
      mx = fmaxf(mx, fmaxf(fabsf(v[r][c].x), fabsf(v[r][c].y)));
    }
  mx = wave_max(mx);
  int ex = 0;
  (void)frexpf(mx, &ex);   // mx < 2^ex
  const int e = (mx > 0.f && mx <= 3.0e38f) ? 8 - ex : 0;
  const float sc = ldexpf(1.f, e);
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int m = l + 64 * c;
      // lanes past the padded row store their zero into its last padding slot (MP > M)
      if (r < nrows) dst[r * MP + min(m, MP - 1)] = make_float2(v[r][c].x * sc, v[r][c].y * sc);
    }
  if constexpr (ZTAIL)
    for (int m = 64 * CB + l; m < MP; m += 64)
      for (int r = 0; r < nrows; ++r) dst[r * MP + m] = make_float2(0.f, 0.f);
  return e;
}

// One staged (scaled) value as FP16 hi + lo, packed (re, im), and the rotated pairs
// (im, -re) that the imaginary part's left operand takes: hi = RN(v), lo = RN(v - hi).
// lo is formed as fma(hi, -1, v) rounded once to FP16 (v - hi is exact in FP32, so this
// is RN(v - hi)); the -1 is hidden from the compiler so that it emits the mixed-precision
// v_fma_mix{lo,hi}_f16 instead of a conversion back to FP32, a subtraction and a second
// conversion.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
struct SplitC {
  unsigned h, l, hr, lr;
};
__device__ __forceinline__ SplitC split_c(float2 v) {
  float m1 = -1.f;
  asm("" : "+v"(m1));
  const half2v hv = {(_Float16)v.x, (_Float16)v.y};
  const half2v lv = {(_Float16)__builtin_fmaf((float)hv.x, m1, v.x), (_Float16)__builtin_fmaf((float)hv.y, m1, v.y)};
  const half2v hr = {hv.y, -hv.x}, lr = {lv.y, -lv.x};
  return {__builtin_bit_cast(unsigned, hv), __builtin_bit_cast(unsigned, lv), __builtin_bit_cast(unsigned, hr),
          __builtin_bit_cast(unsigned, lr)};
}

// D = C + E_t for window row t of the staged rows (scaled by stage_rows_scaled; row stride
// MP, K subarrays) on two v_mfma_f32_32x32x16_f16.  Each operand value is split hi + lo in
// FP16 and the inner index k = 8 (l / 32) + s carries the split products: lane l supplies
// row (and column) l % 32 with
//   A = [a_h b_h a_l b_l]                 (both halves of the wave)
//   B = [a_h -b_h a_h -b_h]  (l < 32),    [a_l -b_l a_l -b_l]  (l >= 32)
// (a = alpha_t, b = beta_t, each entry a packed (re, im) pair), so the products are
// hi.hi + lo.hi in the first half and hi.lo + lo.lo in the second: every term of
// Re E = ar ar + ai ai - br br - bi bi to about 2^-22 of its size, accumulated in FP32.
// The imaginary part takes A with every pair rotated to (im, -re) and the same B.
__device__ __forceinline__ void rank_pass_h(const float2* rows, int MP, int t, int K, const f32x16& cr,
                                            const f32x16& ci, f32x16& dr, f32x16& di) {
  const int i = threadIdx.x & 31;
  const bool h = (threadIdx.x & 32) != 0;
  const float2* xr = rows + t * MP;
  const float2 al = xr[K - 1 + i];
  const float2 bv = xr[i > 0 ? i - 1 : 0];
  const float2 be = i > 0 ? bv : make_float2(0.f, 0.f);
  const SplitC sa = split_c(al), sb = split_c(be);
  constexpr unsigned NEG = 0x80008000u;   // both halves' signs: negation is exact
  const u32x4 A = {sa.h, sb.h, sa.l, sb.l};
  const u32x4 Ai = {sa.hr, sb.hr, sa.lr, sb.lr};
  const unsigned b0 = h ? sa.l : sa.h, b1 = (h ? sb.l : sb.h) ^ NEG;
  const u32x4 B = {b0, b1, b0, b1};
  dr = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(A), as_h8(B), cr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_32x32x16_f16(as_h8(Ai), as_h8(B), ci, 0, 0, 0);
}

// ---- 16 x 16 blocks of the Cholesky trailing update (FP32 matrix cores) ----------------
// In the fused kernel's 16-lane groups (four pixels per wave), lane j of group q holds
// column j of a 16-column slot and row n of a 16-row block in register n: X_q[n][j] at
// (group q, register n).  v_mfma_f32_16x16x1f32 computes four independent 16 x 16
// rank-one products, one per group: block b takes A[i] from lane 16b + i and B[j] from lane
// 16b + j (so row k of the slot, in every group, is both operands as it stands) and keeps
// D_b[i][j] in group i/4, lane j, accumulator element (i%4) + 4b.  With n = 4h + r the two
// layouts differ by swapping the group index with h for each r: a 4 x 4 transpose of
// 16-lane chunks, done by two 32-lane swaps (bit 1 of group and h) and two 16-lane swaps
// (bit 0), and its own inverse.  Register n of the slot layout is element n of the
// accumulator layout.

__device__ __forceinline__ void xpose_groups(f32x16& v) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {   // groups 2,3 of chunk h <-> groups 0,1 of chunk h + 2
      const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[4 * h + r]), __float_as_uint(v[4 * h + 8 + r]),
                                                      false, false);
      v[4 * h + r] = __uint_as_float(s[0]);
      v[4 * h + 8 + r] = __uint_as_float(s[1]);
    }
#pragma unroll
    for (int h = 0; h < 4; h += 2) {   // odd groups of chunk h <-> even groups of chunk h + 1
      const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[4 * h + r]), __float_as_uint(v[4 * h + 4 + r]),
                                                      false, false);
      v[4 * h + r] = __uint_as_float(s[0]);
      v[4 * h + 4 + r] = __uint_as_float(s[1]);
    }
  }
}

// A 4 x 4 transpose of 16-lane chunks over four registers: afterwards x[a] holds, in chunk
// b, what x[b] held in chunk a (two 32-lane swaps, then two 16-lane swaps).
__device__ __forceinline__ void xpose4(float& x0, float& x1, float& x2, float& x3) {
  auto sw32 = [](float& x, float& y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
  };
  auto sw16 = [](float& x, float& y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
  };
  sw32(x0, x2);
  sw32(x1, x3);
  sw16(x0, x1);
  sw16(x2, x3);
}

// The same transpose for doubles (each half of the 64-bit value moved alike).
__device__ __forceinline__ void xpose4d(double& x0, double& x1, double& x2, double& x3) {
  auto lo = [](double v) { return __uint_as_float((unsigned)__builtin_bit_cast(unsigned long long, v)); };
  auto hi = [](double v) { return __uint_as_float((unsigned)(__builtin_bit_cast(unsigned long long, v) >> 32)); };
  auto mk = [](float l, float h) {
    return __builtin_bit_cast(double, (unsigned long long)__float_as_uint(l) |
                                          ((unsigned long long)__float_as_uint(h) << 32));
  };
  float l0 = lo(x0), l1 = lo(x1), l2 = lo(x2), l3 = lo(x3);
  float h0 = hi(x0), h1 = hi(x1), h2 = hi(x2), h3 = hi(x3);
  xpose4(l0, l1, l2, l3);
  xpose4(h0, h1, h2, h3);
  x0 = mk(l0, h0);
  x1 = mk(l1, h1);
  x2 = mk(l2, h2);
  x3 = mk(l3, h3);
}

// D[i][j] -= conj(u_i) v_j for the four groups' 16 x 16 blocks (accumulator layout):
// Re -= ur_i vr_j + ui_i vi_j, Im -= ur_i vi_j - ui_i vr_j, u from the block's rows, v from
// its columns (row k of the respective slots); one of the four MFMAs per call, J = 0..3.
template <int J>
__device__ __forceinline__ void blk_op(f32x16& dr, f32x16& di, float ur, float ui, float vr, float vi) {
  if constexpr (J == 0) dr = __builtin_amdgcn_mfma_f32_16x16x1f32(-ur, vr, dr, 0, 0, 0);
  if constexpr (J == 1) di = __builtin_amdgcn_mfma_f32_16x16x1f32(-ur, vi, di, 0, 0, 0);
  if constexpr (J == 2) dr = __builtin_amdgcn_mfma_f32_16x16x1f32(-ui, vi, dr, 0, 0, 0);
  if constexpr (J == 3) di = __builtin_amdgcn_mfma_f32_16x16x1f32(ui, vr, di, 0, 0, 0);
}

// Number of 16 x 16 blocks (ib, s), 1 <= ib <= s < CPL, still updated on the matrix cores at
// Cholesky step k (those with 16 ib > k).
template <int CPL>
__host__ __device__ constexpr int nlive(int k) {
  int n = 0;
  for (int ib = k / 16 + 1; ib < CPL; ++ib) n += CPL - ib;
  return n;
}

// The window (real sr, imaginary si, MFMA layout) into the 32 x 32 buffer, row i at
// Tb[32 i ..], column j at position 2 (j % 16) + j / 16 of the row: the columns g and
// 16 + g that one lane of a 16-lane group owns sit side by side (one 16-byte read), and
// both the stores (a permutation of each row) and a group's reads (256 contiguous bytes)
// are free of bank conflicts.
__device__ __forceinline__ void window_store(float2* Tb, const f32x16& sr, const f32x16& si) {
  const int c = threadIdx.x & 31, h = (threadIdx.x >> 5) & 1;
  const int pc = ((c & 15) << 1) | (c >> 4);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = 8 * (v / 4) + 4 * h + (v % 4);
    Tb[i * 32 + pc] = make_float2(sr[v], si[v]);
  }
}

}  // namespace cosk
