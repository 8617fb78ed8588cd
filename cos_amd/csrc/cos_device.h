// Device-side helpers for the Capon kernels (gfx950, wave64).
//
// A "group" is the set of lanes that cooperate on one pixel: 16 lanes (one DPP
// row, four pixels per wave) or the whole 64-lane wave.  Cross-lane traffic
// inside a group uses DPP (row_newbcast / row_ror / row_shr) for 16-lane groups
// and v_readlane / wave_ror for 64-lane groups, so no LDS is spent on
// broadcasts.  All loops that index register arrays are unrolled at compile
// time through static_for, which keeps the per-lane matrix columns in VGPRs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace cosk {

template <int N> using ic = std::integral_constant<int, N>;

template <int B, int E, int S = 1, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr ((S > 0 && B < E) || (S < 0 && B > E)) {
    f(ic<B>{});
    static_for<B + S, E, S>(f);
  }
}

template <typename T> struct cplx { T x, y; };

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { return {a, b}; }

// ---- 32-bit DPP / lane primitives ---------------------------------------------------
// mov_dpp leaves lanes whose source is outside the row at 0 (bound_ctrl) and needs
// no zero-initialised destination, unlike update_dpp with an explicit old value.
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_mov_dpp(p.x, CTRL, 0xF, 0xF, true);
  p.y = __builtin_amdgcn_mov_dpp(p.y, CTRL, 0xF, 0xF, true);
  return *reinterpret_cast<double*>(&p);
}
__device__ __forceinline__ float rdlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rdlane(double v, int l) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_readlane(p.x, l);
  p.y = __builtin_amdgcn_readlane(p.y, l);
  return *reinterpret_cast<double*>(&p);
}

// Copies nrows rows of M elements (row stride sstride) from global memory into LDS rows of
// MP elements (zero past M), lanes l0, l0 + nl, ... of the caller taking the columns.  The
// loads of RB rows x CB column chunks are all issued before the first store waits on them:
// a loop that stores each element as it arrives pays one memory latency per element.
template <int RB, int CB>
__device__ __forceinline__ void stage_rows(float2* dst, int MP, const float2* src, size_t sstride, int nrows, int M,
                                           int l0, int nl) {
  for (int t0 = 0; t0 < nrows; t0 += RB)
    for (int m0 = l0; m0 < MP; m0 += nl * CB) {
      float2 v[RB][CB];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int t = t0 + r, m = m0 + nl * c;
          v[r][c] = (t < nrows && m < M) ? src[(size_t)t * sstride + m] : make_float2(0.f, 0.f);
        }
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int t = t0 + r, m = m0 + nl * c;
          if (t < nrows && m < MP) dst[t * MP + m] = v[r][c];
        }
    }
}

// DPP control words (GFX9 encoding).
constexpr int DPP_ROW_SHR = 0x110;      // + n, n in 1..15
constexpr int DPP_ROW_ROR = 0x120;      // + n
constexpr int DPP_WAVE_ROR1 = 0x13C;
constexpr int DPP_ROW_NEWBCAST = 0x150; // + lane in row (gfx90a+)

template <int G> struct Grp;

// 16-lane groups: one DPP row per pixel, four pixels per wave.
template <> struct Grp<16> {
  static constexpr int NG = 4;
  __device__ static int lane() { return threadIdx.x & 15; }
  __device__ static int group() { return (threadIdx.x >> 4) & 3; }
  template <int S, typename T> __device__ static T bcast(T v) { return dpp<DPP_ROW_NEWBCAST + S>(v); }
  template <typename T> __device__ static T ror1(T v) { return dpp<DPP_ROW_ROR + 1>(v); }
  // Sum over the 16 lanes of the row, returned in every lane.
  template <typename T> __device__ static T sum(T v) {
    v += dpp<DPP_ROW_SHR + 1>(v);
    v += dpp<DPP_ROW_SHR + 2>(v);
    v += dpp<DPP_ROW_SHR + 4>(v);
    v += dpp<DPP_ROW_SHR + 8>(v);
    return dpp<DPP_ROW_NEWBCAST + 15>(v);
  }
};

// Whole-wave groups: one pixel per wave.
template <> struct Grp<64> {
  static constexpr int NG = 1;
  __device__ static int lane() { return threadIdx.x & 63; }
  __device__ static int group() { return 0; }
  template <int S, typename T> __device__ static T bcast(T v) { return rdlane(v, S); }
  template <typename T> __device__ static T ror1(T v) { return dpp<DPP_WAVE_ROR1>(v); }
  template <typename T> __device__ static T sum(T v) {
    v += dpp<DPP_ROW_SHR + 1>(v);
    v += dpp<DPP_ROW_SHR + 2>(v);
    v += dpp<DPP_ROW_SHR + 4>(v);
    v += dpp<DPP_ROW_SHR + 8>(v);
    return rdlane(v, 15) + rdlane(v, 31) + rdlane(v, 47) + rdlane(v, 63);
  }
};

template <int G, int S, typename T>
__device__ __forceinline__ cplx<T> bcastc(cplx<T> v) {
  return {Grp<G>::template bcast<S>(v.x), Grp<G>::template bcast<S>(v.y)};
}

// a -= conj(u) * v
template <typename T>
__device__ __forceinline__ void sub_conj_mul(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(-u.x, v.x, a.x);
  a.x = fma(-u.y, v.y, a.x);
  a.y = fma(-u.x, v.y, a.y);
  a.y = fma(u.y, v.x, a.y);
}
// a += u * conj(v)
template <typename T>
__device__ __forceinline__ void add_mul_conj(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(u.x, v.x, a.x);
  a.x = fma(u.y, v.y, a.x);
  a.y = fma(u.y, v.x, a.y);
  a.y = fma(-u.x, v.y, a.y);
}
// a -= u * conj(v)
template <typename T>
__device__ __forceinline__ void sub_mul_conj(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(-u.x, v.x, a.x);
  a.x = fma(-u.y, v.y, a.x);
  a.y = fma(-u.y, v.x, a.y);
  a.y = fma(u.x, v.y, a.y);
}
// a += u * v
template <typename T>
__device__ __forceinline__ void add_mul(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(u.x, v.x, a.x);
  a.x = fma(-u.y, v.y, a.x);
  a.y = fma(u.x, v.y, a.y);
  a.y = fma(u.y, v.x, a.y);
}
// a -= u * v
template <typename T>
__device__ __forceinline__ void sub_mul(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(-u.x, v.x, a.x);
  a.x = fma(u.y, v.y, a.x);
  a.y = fma(-u.x, v.y, a.y);
  a.y = fma(-u.y, v.x, a.y);
}
// a += conj(u) * v
template <typename T>
__device__ __forceinline__ void add_conj_mul(cplx<T>& a, cplx<T> u, cplx<T> v) {
  a.x = fma(u.x, v.x, a.x);
  a.x = fma(u.y, v.y, a.x);
  a.y = fma(u.x, v.y, a.y);
  a.y = fma(-u.y, v.x, a.y);
}

// a -= conj(u_L) * v for one or two accumulators, where u_L is u in lane L of the lane's
// 16-lane row: the broadcast happens inside the FMAs (v_fmac_f32 with a DPP row_newbcast
// source operand) instead of a separate DPP move per component.  NOP adds the two wait
// states a DPP read needs after a VALU write of its source register; without it the
// caller guarantees that u was written at least two instructions earlier.
#define COS_FMAC_BC(D, S, V, NEG) "v_fmac_f32_dpp " D ", " NEG S ", " V " row_newbcast:%[l] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
#define COS_BC1                                                                                               \
  COS_FMAC_BC("%[ax]", "%[ux]", "%[vx]", "-") COS_FMAC_BC("%[ax]", "%[uy]", "%[vy]", "-")                      \
  COS_FMAC_BC("%[ay]", "%[ux]", "%[vy]", "-") COS_FMAC_BC("%[ay]", "%[uy]", "%[vx]", "")
#define COS_BC1_OPS                                                                                           \
  : [ax] "+v"(a.x), [ay] "+v"(a.y) : [ux] "v"(u.x), [uy] "v"(u.y), [vx] "v"(v.x), [vy] "v"(v.y), [l] "n"(L)
#define COS_BC2                                                                                               \
  COS_FMAC_BC("%[ax]", "%[ux]", "%[vx]", "-") COS_FMAC_BC("%[bx]", "%[ux]", "%[wx]", "-")                      \
  COS_FMAC_BC("%[ax]", "%[uy]", "%[vy]", "-") COS_FMAC_BC("%[bx]", "%[uy]", "%[wy]", "-")                      \
  COS_FMAC_BC("%[ay]", "%[ux]", "%[vy]", "-") COS_FMAC_BC("%[by]", "%[ux]", "%[wy]", "-")                      \
  COS_FMAC_BC("%[ay]", "%[uy]", "%[vx]", "") COS_FMAC_BC("%[by]", "%[uy]", "%[wx]", "")
#define COS_BC2_OPS                                                                                           \
  : [ax] "+v"(a.x), [ay] "+v"(a.y), [bx] "+v"(b.x), [by] "+v"(b.y)                                             \
  : [ux] "v"(u.x), [uy] "v"(u.y), [vx] "v"(v.x), [vy] "v"(v.y), [wx] "v"(w.x), [wy] "v"(w.y), [l] "n"(L)
template <int L, bool NOP>
__device__ __forceinline__ void sub_conj_bcast_mul(cplx<float>& a, cplx<float> u, cplx<float> v) {
  if constexpr (NOP)
    asm volatile("s_nop 1\n\t" COS_BC1 COS_BC1_OPS);
  else
    asm volatile(COS_BC1 COS_BC1_OPS);
}
// the same for two accumulators sharing u: a -= conj(u_L) v, b -= conj(u_L) w
template <int L, bool NOP>
__device__ __forceinline__ void sub_conj_bcast_mul2(cplx<float>& a, cplx<float>& b, cplx<float> u, cplx<float> v,
                                                    cplx<float> w) {
  if constexpr (NOP)
    asm volatile("s_nop 1\n\t" COS_BC2 COS_BC2_OPS);
  else
    asm volatile(COS_BC2 COS_BC2_OPS);
}
#undef COS_BC2_OPS
#undef COS_BC2
#undef COS_BC1_OPS
#undef COS_BC1
#undef COS_FMAC_BC

// a -= conj(u_L) v (and b -= conj(u_L) w), u_L = u in lane L of the lane's G-lane group:
// the fused DPP form for FP32 16-lane groups, an explicit broadcast otherwise.  Nothing
// broadcast stays live in registers between calls.  NOP as for sub_conj_bcast_mul: false
// only where u was not written by either of the two preceding instructions (a later
// use of a source that an earlier call already waited for).
template <int G, int L, bool NOP = true, typename T>
__device__ __forceinline__ void bcast_sub_conj_mul(cplx<T>& a, cplx<T> u, cplx<T> v) {
  if constexpr (G == 16 && std::is_same<T, float>::value) {
    sub_conj_bcast_mul<L, NOP>(a, u, v);
  } else {
    sub_conj_mul(a, bcastc<G, L>(u), v);
  }
}
template <int G, int L, bool NOP = true, typename T>
__device__ __forceinline__ void bcast_sub_conj_mul2(cplx<T>& a, cplx<T>& b, cplx<T> u, cplx<T> v, cplx<T> w) {
  if constexpr (G == 16 && std::is_same<T, float>::value) {
    sub_conj_bcast_mul2<L, NOP>(a, b, u, v, w);
  } else {
    bcast_sub_conj_mul<G, L, NOP>(a, u, v);
    bcast_sub_conj_mul<G, L, NOP>(b, u, w);
  }
}

// sin(pi x) and cos(pi x) in the precision of T (the FP64 kernels take steering phases in
// double, so their power and weights carry no FP32 phase rounding)
template <typename T> __device__ __forceinline__ void sincospi_t(T x, T* s, T* c);
template <> __device__ __forceinline__ void sincospi_t<float>(float x, float* s, float* c) { sincospif(x, s, c); }
template <> __device__ __forceinline__ void sincospi_t<double>(double x, double* s, double* c) { sincospi(x, s, c); }

template <typename T> __device__ __forceinline__ T rsqrt_t(T v);
// v_rsq_f32 (1 ulp); the FP32 path's accuracy budget is set by the covariance sums.
template <> __device__ __forceinline__ float rsqrt_t<float>(float v) { return __builtin_amdgcn_rsqf(v); }
template <> __device__ __forceinline__ double rsqrt_t<double>(double v) { return 1.0 / sqrt(v); }

}  // namespace cosk
