// Component kernels around the fused Capon kernel: fallback compaction, the batched
// general solver behind Solver::solve, the beamspace covariance behind
// BuildR::getR(Nb > 0), and the reference's matrix-beamspace output path.
#include "cos_device.h"
#include "components.h"
#include "capon_kernels.h"

namespace cosk {

__global__ void fill_i32_kernel(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Pixels with any flag bit go to the FP64 list (order is irrelevant: pixels are
// independent and each writes only its own outputs); with a row-L table, to the list of
// the bucket of their row's L (list b at list + b * lstride, its length in count[b]), so
// each bucket runs on the lane shape of its L.  Each thread tests sixteen pixels (one
// 16-byte load where they are in range and aligned), and per bucket a block-wide prefix
// of the threads' counts, in pixel order, and one atomic per block reserve the block's
// entries: neighbouring pixels stay neighbours in the list (one atomic per wave and
// bucket, over four pixels per lane, had the whole grid contend for the counters: 0.20 ms
// for the expanding-aperture frame's 1,047,552 flags).
__global__ __launch_bounds__(256) void compact_flags_kernel(const uint8_t* flags, int n, int* list, int* count,
                                                            const int* rowL, int nzp, int Y, RowBuckets bk, int nb,
                                                            int lstride) {
  __shared__ int wsum[4];
  __shared__ int bbase;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int i0 = 16 * (blockIdx.x * 256 + t);
  uint32_t f[4] = {0u, 0u, 0u, 0u};
  if (i0 + 15 < n && ((uintptr_t)(flags + i0) & 15) == 0) {
    const uint4 v = *reinterpret_cast<const uint4*>(flags + i0);
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
    f[3] = v.w;
  } else {
    for (int u = 0; u < 16; ++u)
      if (i0 + u < n) f[u >> 2] |= (uint32_t)flags[i0 + u] << (8 * (u & 3));
  }
  // bucket + 1 of each flagged pixel, four bits each (0: not flagged)
  unsigned long long code = 0;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    if ((f[u >> 2] >> (8 * (u & 3))) & 0xffu) {
      int b = 0;
      if (rowL) {
        int L = rowL[(i0 + u) % nzp + Y];
        L = L < 0 ? 0 : L > 64 ? 64 : L;
        b = bk.b[L];
      }
      code |= (unsigned long long)(b + 1) << (4 * u);
    }
  }
  for (int bb = 0; bb < nb; ++bb) {
    int c = 0;
#pragma unroll
    for (int u = 0; u < 16; ++u) c += ((code >> (4 * u)) & 15u) == (unsigned long long)(bb + 1);
    int inc = c;   // inclusive prefix over the wave's lanes
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    int wbase = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      wbase += w < wv ? wsum[w] : 0;
      tot += wsum[w];
    }
    if (t == 0) bbase = tot ? atomicAdd(count + bb, tot) : 0;
    __syncthreads();
    int pos = bbase + wbase + inc - c;
    if (c)
      for (int u = 0; u < 16; ++u)
        if (((code >> (4 * u)) & 15u) == (unsigned long long)(bb + 1)) list[(size_t)bb * lstride + pos++] = i0 + u;
    __syncthreads();   // wsum and bbase are reused by the next bucket
  }
}

// Row lists of a table-driven frame: one block of 1024 threads takes the output rows in
// chunks of 1024; each row goes to the list of its bucket (the lane shape of its L), in
// increasing row order, by a ballot prefix within each wave and a sum over the waves.
__global__ __launch_bounds__(1024) void plan_rows_kernel(const int* rowL, int nzp, int Y, RowBuckets bk, int nb,
                                                         int* lists, int* counts) {
  constexpr int NW = 1024 / 64, NB = 8;
  __shared__ int wtot[NW][NB];
  __shared__ int base[NB];
  const int t = threadIdx.x, wv = t >> 6, ln = t & 63;
  if (t < NB) base[t] = 0;
  __syncthreads();
  for (int j0 = 0; j0 < nzp; j0 += 1024) {
    const int j = j0 + t;
    int b = -1;
    if (j < nzp) {
      int L = rowL[j + Y];
      L = L < 0 ? 0 : L > 64 ? 64 : L;
      b = bk.b[L];
    }
    unsigned long long mine = 0;
    for (int i = 0; i < nb; ++i) {
      const unsigned long long m = __ballot(b == i);
      if (b == i) mine = m;
      if (ln == 0) wtot[wv][i] = __popcll(m);
    }
    __syncthreads();
    if (b >= 0) {
      int off = base[b];
      for (int w = 0; w < wv; ++w) off += wtot[w][b];
      off += __popcll(mine & ((1ull << ln) - 1ull));
      lists[b * nzp + off] = j;
    }
    __syncthreads();
    if (t < nb) {
      int n = 0;
      for (int w = 0; w < NW; ++w) n += wtot[w][t];
      base[t] += n;
    }
    __syncthreads();
  }
  if (t < nb) counts[t] = base[t];
}

hipError_t plan_rows(const int* rowL, int nzp, int Y, RowBuckets bk, int nb, int* lists, int* counts, hipStream_t s) {
  if (nb < 1 || nb > 8 || nzp < 1) return hipErrorInvalidValue;
  for (int L = 0; L <= 64; ++L)
    if (bk.b[L] >= nb) return hipErrorInvalidValue;
  hipLaunchKernelGGL(plan_rows_kernel, dim3(1), dim3(1024), 0, s, rowL, nzp, Y, bk, nb, lists, counts);
  return hipGetLastError();
}

hipError_t fill_i32(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t compact_flags(const uint8_t* flags, int n, int* list, int* count, hipStream_t s) {
  return compact_flags_rows(flags, n, list, count, nullptr, 1, 0, RowBuckets{}, 1, 0, s);
}

hipError_t compact_flags_rows(const uint8_t* flags, int n, int* list, int* count, const int* rowL, int nzp, int Y,
                              RowBuckets bk, int nb, int lstride, hipStream_t s) {
  if (nb < 1 || nb > 8 || (rowL && nzp < 1)) return hipErrorInvalidValue;
  if (rowL)
    for (int L = 0; L <= 64; ++L)
      if (bk.b[L] >= nb) return hipErrorInvalidValue;
  hipError_t e = fill_i32(count, nb, 0, s);
  if (e != hipSuccess || n <= 0) return e;
  hipLaunchKernelGGL(compact_flags_kernel, dim3((n + 4095) / 4096), dim3(256), 0, s, flags, n, list, count, rowL,
                     nzp, Y, bk, nb, lstride);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Batched general complex solve A x = b, A column-major (A[i + N*j] = A_ij), one wave
// per system, the augmented matrix [N][N+1] in LDS, Gaussian elimination with partial
// pivoting on |re| + |im| (the pivot measure of RealTimeCapon/operations.h:183-207; the
// first row of largest measure wins, as a sequential scan picks it).  Each step: the
// pivot by a wave-wide arg-max, the row swap and the multipliers across lanes, then the
// trailing update with lanes along columns, so a lane's column stays in consecutive LDS
// words and the multiplier of a row is one broadcast read.  The back substitution runs
// column by column with lanes over the rows above.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(64) void solve_lu_kernel(const float2* A, const float2* b, float2* x,
                                                      int N, int batch) {
  extern __shared__ float2 sm[];
  const int sys = blockIdx.x;
  if (sys >= batch) return;
  const int lane = threadIdx.x;
  const int NP = N + 1;           // row stride of the augmented matrix [N][N+1]
  float2* M = sm;
  const float2* As = A + (size_t)sys * N * N;
  for (int e = lane; e < N * N; e += 64) {
    const int i = e % N, jj = e / N;
    M[i * NP + jj] = As[e];
  }
  for (int i = lane; i < N; i += 64) M[i * NP + N] = b[(size_t)sys * N + i];
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    float bv = -1.f;
    int bi = k;
    for (int i = k + lane; i < N; i += 64) {
      const float2 v = M[i * NP + k];
      const float m = fabsf(v.x) + fabsf(v.y);
      if (m > bv) { bv = m; bi = i; }
    }
    // (measure descending, row ascending) is a total order: every lane ends on the same row
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o);
      const int oi = __shfl_xor(bi, o);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    const int pr = __builtin_amdgcn_readfirstlane(bi);
    if (pr != k) {
      for (int jj = k + lane; jj <= N; jj += 64) {
        const float2 t = M[k * NP + jj];
        M[k * NP + jj] = M[pr * NP + jj];
        M[pr * NP + jj] = t;
      }
      __syncthreads();
    }
    const float2 pv = M[k * NP + k];
    const float den = pv.x * pv.x + pv.y * pv.y;
    const float2 ipv = make_float2(pv.x / den, -pv.y / den);
    // multipliers f_i = a_ik / a_kk in place of column k (not read again)
    for (int i = k + 1 + lane; i < N; i += 64) M[i * NP + k] = cmul(M[i * NP + k], ipv);
    __syncthreads();
    for (int jj = k + 1 + lane; jj <= N; jj += 64) {
      const float2 u = M[k * NP + jj];
      float2* col = M + jj;
      int i = k + 1;
      for (; i + 1 < N; i += 2) {
        const float2 f0 = M[i * NP + k], f1 = M[(i + 1) * NP + k];
        float2 v0 = col[i * NP], v1 = col[(i + 1) * NP];
        v0.x -= f0.x * u.x - f0.y * u.y;
        v0.y -= f0.x * u.y + f0.y * u.x;
        v1.x -= f1.x * u.x - f1.y * u.y;
        v1.y -= f1.x * u.y + f1.y * u.x;
        col[i * NP] = v0;
        col[(i + 1) * NP] = v1;
      }
      if (i < N) {
        const float2 f0 = M[i * NP + k];
        float2 v0 = col[i * NP];
        v0.x -= f0.x * u.x - f0.y * u.y;
        v0.y -= f0.x * u.y + f0.y * u.x;
        col[i * NP] = v0;
      }
    }
    __syncthreads();
  }
  // U x = c: x_i = c_i / u_ii, then c_r -= u_ri x_i for the rows r < i
  for (int i = N - 1; i >= 0; --i) {
    const float2 s = M[i * NP + N], d = M[i * NP + i];
    const float den = d.x * d.x + d.y * d.y;
    const float2 r = make_float2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den);
    __syncthreads();
    if (lane == 0) M[i * NP + N] = r;
    for (int rr = lane; rr < i; rr += 64) {
      const float2 u = M[rr * NP + i];
      float2 c = M[rr * NP + N];
      c.x -= u.x * r.x - u.y * r.y;
      c.y -= u.x * r.y + u.y * r.x;
      M[rr * NP + N] = c;
    }
    __syncthreads();
  }
  for (int i = lane; i < N; i += 64) x[(size_t)sys * N + i] = M[i * NP + N];
}

// ---------------------------------------------------------------------------------
// The same elimination with the rows in registers, for N <= 64: a group of R lanes holds
// one system, lane i its row i of the augmented matrix [A | b] (NC = R + 1 complex values,
// statically indexed; the right-hand side in the last slot), 64/R systems per wave.  A row
// exchange is a relabelling: each lane keeps the position its row has reached, and the
// pivot of step k is the unused row of largest |re|+|im| in column k, ties to the lowest
// position (what the sequential scan of the LDS kernel picks).  The pivot row goes through
// LDS to its group as one broadcast read per column pair (measured faster than two lane
// permutes per value: 2.88 against 3.49 ms for 523,264 systems of N = 32); x_k of the
// back substitution comes from the pivot lane by a lane permute.
// ---------------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u(unsigned v) {
  return (unsigned)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// max (MIN = false) or min over each aligned group of R lanes; every lane must be active
template <int R, bool MIN>
__device__ __forceinline__ unsigned group_reduce(unsigned v) {
  auto op = [](unsigned a, unsigned b2) { return MIN ? min(a, b2) : max(a, b2); };
  v = op(v, dpp_u<0xB1>(v));                     // quad_perm [1,0,3,2]
  if (R > 2) v = op(v, dpp_u<0x4E>(v));          // quad_perm [2,3,0,1]
  if (R > 4) v = op(v, dpp_u<0x141>(v));         // row_half_mirror
  if (R > 8) v = op(v, dpp_u<0x140>(v));         // row_mirror
  if (R > 16) v = op(v, (unsigned)__shfl_xor((int)v, 16));
  if (R > 32) v = op(v, (unsigned)__shfl_xor((int)v, 32));
  return v;
}

template <int R>
__global__ __launch_bounds__(256) void solve_rows_kernel(const float2* __restrict__ A, const float2* __restrict__ b,
                                                         float2* __restrict__ x, int N, int batch) {
  constexpr int NC = R + 1, S = 64 / R, NCP = (NC + 1) & ~1;
  __shared__ __align__(16) float2 prow[4][S][NCP];   // each group's pivot row
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane / R, i = lane % R;
  const long sys = ((long)blockIdx.x * 4 + wv) * S + g;
  const bool live = sys < batch && i < N;
  float2 a[NC];
  {
    const float2* As = A + (live ? sys : 0) * (long)N * N + i;
#pragma unroll
    for (int j = 0; j < NC - 1; ++j) a[j] = (live && j < N) ? As[(long)N * j] : make_float2(0.f, 0.f);
    a[NC - 1] = live ? b[sys * N + i] : make_float2(0.f, 0.f);
  }
  int pos = i;        // position of this lane's row after the exchanges so far
  bool act = live;    // not yet used as a pivot
  unsigned plk[(NC + 2) / 4] = {};   // the pivot lane of each step, four to a register
#pragma unroll
  for (int k = 0; k < NC - 1; ++k) {
    if (k >= N) continue;
    const float m = fabsf(a[k].x) + fabsf(a[k].y);
    // |re|+|im| >= 0 orders as its bit pattern; 0 marks rows out of the search (and NaN)
    const unsigned key = (act && m == m) ? __float_as_uint(m) + 1u : 0u;
    const unsigned top = group_reduce<R, false>(key);
    // among the rows at the maximum (all unused rows when none holds a number), the lowest
    // position, with its lane in the low six bits
    const unsigned sel = group_reduce<R, true>(act && key == top ? ((unsigned)pos << 6) | (unsigned)lane : ~0u);
    if (sel != ~0u) {   // the same in every lane of a group (a group past the batch has none)
      const int pl = (int)(sel & 63u), pp = (int)(sel >> 6);
      plk[k / 4] |= (unsigned)pl << (8 * (k % 4));
      if (lane == pl) {
        pos = k;
        act = false;
      } else if (act && pos == k) {
        pos = pp;   // the row at position k takes the pivot row's place
      }
      // the pivot lane publishes its row (LDS operations of a wave complete in order);
      // rows out of the elimination (used pivots, padding) take f = 0
      float2* pr = prow[wv][g];
      if (lane == pl) {
#pragma unroll
        for (int j = k; j < NC; ++j) pr[j] = a[j];
      }
      __builtin_amdgcn_wave_barrier();
      const float2 u = pr[k];
      const float rden = __builtin_amdgcn_rcpf(u.x * u.x + u.y * u.y);   // 1 ulp
      float2 f = cmul(a[k], make_float2(u.x * rden, -u.y * rden));
      if (!act) f = make_float2(0.f, 0.f);
      // a_j -= f u_j as four FMAs; columns N .. R-1 of a short system are zero and stay zero
#pragma unroll
      for (int j = k + 1; j < NC; ++j) {
        const float2 v = pr[j];
        a[j].x = fmaf(-f.x, v.x, fmaf(f.y, v.y, a[j].x));
        a[j].y = fmaf(-f.x, v.y, fmaf(-f.y, v.x, a[j].y));
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // back substitution: step k's pivot lane holds U's row k and gives x_k, the rows of the
  // earlier positions update (every lane forms a candidate; the group takes the pivot's)
  float2 xv = make_float2(0.f, 0.f);
#pragma unroll
  for (int k = NC - 2; k >= 0; --k) {
    if (k < N) {
      const float2 s2 = a[NC - 1], d = a[k];
      const float rden = __builtin_amdgcn_rcpf(d.x * d.x + d.y * d.y);
      const float2 xk = make_float2((s2.x * d.x + s2.y * d.y) * rden, (s2.y * d.x - s2.x * d.y) * rden);
      const int pl = (int)((plk[k / 4] >> (8 * (k % 4))) & 63u);
      float2 r = make_float2(__shfl(xk.x, pl), __shfl(xk.y, pl));
      if (lane == pl) xv = r;
      if (pos >= k) r = make_float2(0.f, 0.f);
      a[NC - 1].x = fmaf(-a[k].x, r.x, fmaf(a[k].y, r.y, a[NC - 1].x));
      a[NC - 1].y = fmaf(-a[k].x, r.y, fmaf(-a[k].y, r.x, a[NC - 1].y));
    }
  }
  if (live) x[sys * N + pos] = xv;
}

hipError_t solve_lu(const float2* A, const float2* b, float2* x, int N, int batch, hipStream_t s) {
  if (batch <= 0) return hipSuccess;
  if (N < 1 || N > 72) return hipErrorInvalidValue;
  auto rows = [&](auto kern, int R) {
    const long per = 4L * (64 / R);
    hipLaunchKernelGGL(kern, dim3((unsigned)((batch + per - 1) / per)), dim3(256), 0, s, A, b, x, N, batch);
    return hipGetLastError();
  };
  if (N <= 8) return rows(solve_rows_kernel<8>, 8);
  if (N <= 16) return rows(solve_rows_kernel<16>, 16);
  if (N <= 32) return rows(solve_rows_kernel<32>, 32);
  if (N <= 64) return rows(solve_rows_kernel<64>, 64);
  size_t lds = (size_t)N * (N + 1) * sizeof(float2);
  hipLaunchKernelGGL(solve_lu_kernel, dim3(batch), dim3(64), lds, s, A, b, x, N, batch);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// BuildR with Nb > 0: x_bs [Nx][Nz][K][Nb]; R[r][c] over k and the window, loading
// d*trace/Nb, stored conjugated like buildR_full_kernel (buildR_kernel.cu:169-251).
// One thread per (pixel, r, c); the trace is recomputed per thread (Nb is small).
// ---------------------------------------------------------------------------------
__global__ void buildr_bs_kernel(const float2* xbs, float2* R, float d, int Nb, int K, int Y,
                                 int Nx, int Nz) {
  const int Nzp = Nz - 2 * Y;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = (long)Nx * Nzp * Nb * Nb;
  if (tid >= tot) return;
  const int e = (int)(tid % (Nb * Nb));
  const long p = tid / (Nb * Nb);
  const int r = e / Nb, c = e % Nb;
  const int ix = (int)(p / Nzp), j = (int)(p % Nzp);
  double sr = 0, si = 0, tr = 0;
  for (int t = 0; t < 2 * Y + 1; ++t) {
    const float2* row = xbs + ((size_t)ix * Nz + j + t) * (size_t)K * Nb;
    for (int k = 0; k < K; ++k) {
      float2 u = row[k * Nb + r], v = row[k * Nb + c];
      sr += (double)u.x * v.x + (double)u.y * v.y;
      si += (double)u.y * v.x - (double)u.x * v.y;
      for (int l = 0; l < Nb; ++l) {
        float2 w = row[k * Nb + l];
        tr += (double)w.x * w.x + (double)w.y * w.y;
      }
    }
  }
  if (r == c) sr += (double)d * tr / Nb;
  R[p * Nb * Nb + (size_t)r * Nb + c] = make_float2((float)sr, (float)-si);
}

hipError_t buildr_beamspace(const float2* xbs, float2* R, float d, int Nb, int K, int Y, int Nx,
                            int Nz, hipStream_t s) {
  long tot = (long)Nx * (Nz - 2 * Y) * Nb * Nb;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(buildr_bs_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, xbs, R,
                     d, Nb, K, Y, Nx, Nz);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Reference default beamspace path (RealTimeCapon/Capon.cpp:285-402) from the
// materialised conj-stored S_L: R_bs = B S_L B^H, a_bs = B 1, y_bs = R_bs^-1 a_bs,
// y = B^H y_bs, then the element-space output stage (capon_kernel.cu:338-427):
// w = y / sum(y), z = (1/K) sum_l conj(w_l) sum_k x_c[k+l].  P = 1/(norm a_bs^H y_bs).
// One 64-thread block per pixel; B is [Nb][L] row-major in global memory.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void beamspace_matrix_kernel(const float2* Rg, const float2* B,
                                                              const float2* x, int L, int Nb, int M,
                                                              int Y, int Nx, int Nz, float2* w,
                                                              float2* z, float* P) {
  extern __shared__ float2 sm[];
  const int Nzp = Nz - 2 * Y;
  const int p = blockIdx.x;
  if (p >= Nx * Nzp) return;
  float2* S = sm;                 // [L][L] S_L[r][c]
  float2* Tm = S + L * L;         // [L][Nb] S_L B^H
  float2* Rb = Tm + L * Nb;       // [Nb][Nb]
  float2* yb = Rb + Nb * Nb;      // [Nb]
  float2* ab = yb + Nb;           // [Nb]
  float2* y = ab + Nb;            // [L]
  const float2* Rp = Rg + (size_t)p * L * L;
  for (int e = threadIdx.x; e < L * L; e += 64) {
    int r = e / L, c = e % L;
    float2 v = Rp[e];  // conj(S_L[r][c])
    S[r * L + c] = make_float2(v.x, -v.y);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < L * Nb; e += 64) {
    int r = e / Nb, bb = e % Nb;
    double sx = 0, sy = 0;
    for (int c = 0; c < L; ++c) {  // S[r][c] conj(B[bb][c])
      float2 s = S[r * L + c], v = B[bb * L + c];
      sx += (double)s.x * v.x + (double)s.y * v.y;
      sy += (double)s.y * v.x - (double)s.x * v.y;
    }
    Tm[r * Nb + bb] = make_float2((float)sx, (float)sy);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < Nb * Nb; e += 64) {
    int b1 = e / Nb, b2 = e % Nb;
    double sx = 0, sy = 0;
    for (int r = 0; r < L; ++r) {  // B[b1][r] T[r][b2]
      float2 u = B[b1 * L + r], v = Tm[r * Nb + b2];
      sx += (double)u.x * v.x - (double)u.y * v.y;
      sy += (double)u.x * v.y + (double)u.y * v.x;
    }
    Rb[b1 * Nb + b2] = make_float2((float)sx, (float)sy);
  }
  for (int bb = threadIdx.x; bb < Nb; bb += 64) {
    double sx = 0, sy = 0;
    for (int l = 0; l < L; ++l) { sx += B[bb * L + l].x; sy += B[bb * L + l].y; }
    ab[bb] = make_float2((float)sx, (float)sy);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // Cholesky U^H U = Rb in place (upper triangle), then U^H v = ab, U yb = v; double precision
    // accumulation, Nb is small.
    for (int i = 0; i < Nb; ++i) {
      double dr = Rb[i * Nb + i].x;
      for (int k = 0; k < i; ++k) {
        float2 u = Rb[k * Nb + i];
        dr -= (double)u.x * u.x + (double)u.y * u.y;
      }
      double di = sqrt(dr > 0 ? dr : 1e-30);
      Rb[i * Nb + i] = make_float2((float)di, 0.f);
      for (int c = i + 1; c < Nb; ++c) {
        double sx = Rb[i * Nb + c].x, sy = Rb[i * Nb + c].y;
        for (int k = 0; k < i; ++k) {
          float2 u = Rb[k * Nb + i], v = Rb[k * Nb + c];  // conj(u) v
          sx -= (double)u.x * v.x + (double)u.y * v.y;
          sy -= (double)u.x * v.y - (double)u.y * v.x;
        }
        Rb[i * Nb + c] = make_float2((float)(sx / di), (float)(sy / di));
      }
    }
    for (int i = 0; i < Nb; ++i) {
      double sx = ab[i].x, sy = ab[i].y;
      for (int k = 0; k < i; ++k) {
        float2 u = Rb[k * Nb + i], v = yb[k];
        sx -= (double)u.x * v.x + (double)u.y * v.y;
        sy -= (double)u.x * v.y - (double)u.y * v.x;
      }
      double dd = Rb[i * Nb + i].x;
      yb[i] = make_float2((float)(sx / dd), (float)(sy / dd));
    }
    for (int i = Nb - 1; i >= 0; --i) {
      double sx = yb[i].x, sy = yb[i].y;
      for (int k = i + 1; k < Nb; ++k) {
        float2 u = Rb[i * Nb + k], v = yb[k];
        sx -= (double)u.x * v.x - (double)u.y * v.y;
        sy -= (double)u.x * v.y + (double)u.y * v.x;
      }
      double dd = Rb[i * Nb + i].x;
      yb[i] = make_float2((float)(sx / dd), (float)(sy / dd));
    }
    double qx = 0;
    for (int i = 0; i < Nb; ++i) qx += (double)ab[i].x * yb[i].x + (double)ab[i].y * yb[i].y;
    const int K = M - L + 1;
    if (P) P[p] = (float)(1.0 / (qx * (double)K * (2 * Y + 1)));
  }
  __syncthreads();
  for (int l = threadIdx.x; l < L; l += 64) {  // y = B^H yb
    double sx = 0, sy = 0;
    for (int bb = 0; bb < Nb; ++bb) {
      float2 u = B[bb * L + l], v = yb[bb];  // conj(u) v
      sx += (double)u.x * v.x + (double)u.y * v.y;
      sy += (double)u.x * v.y - (double)u.y * v.x;
    }
    y[l] = make_float2((float)sx, (float)sy);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0, sy = 0;
    for (int l = 0; l < L; ++l) { sx += y[l].x; sy += y[l].y; }
    const double den = sx * sx + sy * sy;
    const int ix = p / Nzp, j = p % Nzp, K = M - L + 1;
    const float2* xc = x + ((size_t)ix * Nz + j + Y) * M;
    double zx = 0, zy = 0;
    for (int l = 0; l < L; ++l) {
      double wx = (y[l].x * sx + y[l].y * sy) / den, wy = (y[l].y * sx - y[l].x * sy) / den;
      if (w) w[(size_t)p * L + l] = make_float2((float)wx, (float)wy);
      double gx = 0, gy = 0;
      for (int k = 0; k < K; ++k) { gx += xc[k + l].x; gy += xc[k + l].y; }
      zx += wx * gx + wy * gy;
      zy += wx * gy - wy * gx;
    }
    if (z) z[p] = make_float2((float)(zx / K), (float)(zy / K));
  }
}

// ---------------------------------------------------------------------------------
// Capon in the subspace of a user matrix V [Nb][L] (framework/beamformer/capon/
// getCaponAllC.pyx:230,346-447, loading in the subspace, its default): from the
// conj-stored unnormalised S of one pixel, R = V (S/norm) V^H, loading d tr(R)/Nb,
// a = V 1, g = V g_es with g_es[l] = (1/K) sum_k x_c[k+l], y = R_L^-1 a,
// w = y/(a^H y) (Nb entries), z = w^H g, P = 1/(a^H y).  One 64-thread block per pixel,
// FP64 accumulation (Nb and L are small).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void subspace_capon_kernel(const float2* Rg, const float2* V,
                                                            const float2* x, int L, int Nb, int M,
                                                            int Mstride, int off, int Y, int Nx, int Nz,
                                                            float d, float2* w, float2* z, float* P) {
  extern __shared__ double2 smd[];
  const int Nzp = Nz - 2 * Y, K = M - L + 1;
  const int p = blockIdx.x;
  if (p >= Nx * Nzp) return;
  double2* Tm = smd;              // [L][Nb] (S/norm) V^H

  double2* Rb = Tm + L * Nb;      // [Nb][Nb]
  double2* ab = Rb + Nb * Nb;     // [Nb] V 1
  double2* gb = ab + Nb;          // [Nb] V g_es
  double2* yb = gb + Nb;          // [Nb]
  double2* ge = yb + Nb;          // [L] g_es
  float2* S = reinterpret_cast<float2*>(ge + L);   // [L][L] S[r][c] (unnormalised)
  const double inorm = 1.0 / ((double)K * (2 * Y + 1));
  const float2* Rp = Rg + (size_t)p * L * L;
  for (int e = threadIdx.x; e < L * L; e += 64) {
    const float2 v = Rp[e];   // conj(S[r][c])
    S[e] = make_float2(v.x, -v.y);
  }
  const int ix = p / Nzp, j = p % Nzp;
  const float2* xc = x + ((size_t)ix * Nz + j + Y) * Mstride + off;
  for (int l = threadIdx.x; l < L; l += 64) {
    double sx = 0, sy = 0;
    for (int k = 0; k < K; ++k) { sx += xc[k + l].x; sy += xc[k + l].y; }
    ge[l] = make_double2(sx / K, sy / K);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < L * Nb; e += 64) {
    const int r = e / Nb, bb = e % Nb;
    double sx = 0, sy = 0;
    for (int c = 0; c < L; ++c) {   // S[r][c] conj(V[bb][c])
      const float2 sv = S[r * L + c];
      const float2 v = V[bb * L + c];
      sx += (double)sv.x * v.x + (double)sv.y * v.y;
      sy += (double)sv.y * v.x - (double)sv.x * v.y;
    }
    Tm[e] = make_double2(sx * inorm, sy * inorm);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < Nb * Nb; e += 64) {
    const int b1 = e / Nb, b2 = e % Nb;
    double sx = 0, sy = 0;
    for (int r = 0; r < L; ++r) {   // V[b1][r] T[r][b2]
      const float2 u = V[b1 * L + r];
      const double2 v = Tm[r * Nb + b2];
      sx += u.x * v.x - u.y * v.y;
      sy += u.x * v.y + u.y * v.x;
    }
    Rb[e] = make_double2(sx, sy);
  }
  for (int bb = threadIdx.x; bb < Nb; bb += 64) {
    double ax = 0, ay = 0, gx = 0, gy = 0;
    for (int l = 0; l < L; ++l) {
      const float2 v = V[bb * L + l];
      ax += v.x;
      ay += v.y;
      gx += v.x * ge[l].x - v.y * ge[l].y;
      gy += v.x * ge[l].y + v.y * ge[l].x;
    }
    ab[bb] = make_double2(ax, ay);
    gb[bb] = make_double2(gx, gy);
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tr = 0;
  for (int i = 0; i < Nb; ++i) tr += Rb[i * Nb + i].x;
  const double load = (double)d * tr / Nb;
  // Cholesky U^H U = Rb + load I (upper triangle in place), U^H v = a, U y = v
  for (int i = 0; i < Nb; ++i) {
    double dr = Rb[i * Nb + i].x + load;
    for (int k = 0; k < i; ++k) dr -= Rb[k * Nb + i].x * Rb[k * Nb + i].x + Rb[k * Nb + i].y * Rb[k * Nb + i].y;
    const double di = sqrt(dr > 1e-300 ? dr : 1e-300);
    Rb[i * Nb + i] = make_double2(di, 0.0);
    for (int c = i + 1; c < Nb; ++c) {
      double sx = Rb[i * Nb + c].x, sy = Rb[i * Nb + c].y;
      for (int k = 0; k < i; ++k) {   // conj(U[k][i]) U[k][c]
        const double2 u = Rb[k * Nb + i], v = Rb[k * Nb + c];
        sx -= u.x * v.x + u.y * v.y;
        sy -= u.x * v.y - u.y * v.x;
      }
      Rb[i * Nb + c] = make_double2(sx / di, sy / di);
    }
  }
  for (int i = 0; i < Nb; ++i) {
    double sx = ab[i].x, sy = ab[i].y;
    for (int k = 0; k < i; ++k) {
      const double2 u = Rb[k * Nb + i], v = yb[k];
      sx -= u.x * v.x + u.y * v.y;
      sy -= u.x * v.y - u.y * v.x;
    }
    yb[i] = make_double2(sx / Rb[i * Nb + i].x, sy / Rb[i * Nb + i].x);
  }
  for (int i = Nb - 1; i >= 0; --i) {
    double sx = yb[i].x, sy = yb[i].y;
    for (int k = i + 1; k < Nb; ++k) {
      const double2 u = Rb[i * Nb + k], v = yb[k];
      sx -= u.x * v.x - u.y * v.y;
      sy -= u.x * v.y + u.y * v.x;
    }
    yb[i] = make_double2(sx / Rb[i * Nb + i].x, sy / Rb[i * Nb + i].x);
  }
  double qx = 0, qy = 0;   // a^H y
  for (int i = 0; i < Nb; ++i) {
    qx += ab[i].x * yb[i].x + ab[i].y * yb[i].y;
    qy += ab[i].x * yb[i].y - ab[i].y * yb[i].x;
  }
  const double den = qx * qx + qy * qy;
  double zx = 0, zy = 0;
  for (int i = 0; i < Nb; ++i) {
    const double wx = (yb[i].x * qx + yb[i].y * qy) / den, wy = (yb[i].y * qx - yb[i].x * qy) / den;
    if (w) w[(size_t)p * Nb + i] = make_float2((float)wx, (float)wy);
    zx += wx * gb[i].x + wy * gb[i].y;   // conj(w) g
    zy += wx * gb[i].y - wy * gb[i].x;
  }
  if (z) z[p] = make_float2((float)zx, (float)zy);
  if (P) P[p] = (float)(1.0 / qx);
}

hipError_t subspace_capon(const float2* R, const float2* V, const float2* x, int L, int Nb, int M, int Mstride,
                          int off, int Y, int Nx, int Nz, float d, float2* w, float2* z, float* P, hipStream_t s) {
  const int npx = Nx * (Nz - 2 * Y);
  if (npx <= 0) return hipSuccess;
  const size_t lds = ((size_t)L * Nb + (size_t)Nb * Nb + 3 * (size_t)Nb + L) * sizeof(double2) +
                     (size_t)L * L * sizeof(float2);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(subspace_capon_kernel, dim3(npx), dim3(64), lds, s, R, V, x, L, Nb, M, Mstride, off, Y,
                     Nx, Nz, d, w, z, P);
  return hipGetLastError();
}

hipError_t beamspace_matrix(const float2* R, const float2* B, const float2* x, int L, int Nb, int M,
                            int Y, int Nx, int Nz, float2* w, float2* z, float* P, hipStream_t s) {
  const int npx = Nx * (Nz - 2 * Y);
  if (npx <= 0) return hipSuccess;
  size_t lds = ((size_t)L * L + (size_t)L * Nb + (size_t)Nb * Nb + 2 * Nb + L) * sizeof(float2);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(beamspace_matrix_kernel, dim3(npx), dim3(64), lds, s, R, B, x, L, Nb, M, Y, Nx,
                     Nz, w, z, P);
  return hipGetLastError();
}

}  // namespace cosk
