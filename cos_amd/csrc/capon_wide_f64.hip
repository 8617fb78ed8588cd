// Per-pixel FP64 Capon for wide subarrays (L up to 64): one wave per pixel with the
// matrix in LDS (packed upper triangle, column-major) and plain loops.  It serves the
// all-FP64 run and the FP64 recomputation of flagged pixels when a frame has L > 32.
// A 64 x 64 complex double matrix does not fit one lane group's registers (it would
// need 256 VGPRs per lane for the matrix alone), so the factorisation works in LDS:
// lane c owns column c, row k of U is published in LDS and read back as a broadcast.
//
// Semantics (as in capon_fused_impl.h; paths relative to the reference root):
//   covariance S[r][c] = sum_t sum_{k<K} x_t[k+r] conj(x_t[k+c]) in FP64: the first row
//              by direct sums, the rest by the displacement recursion along diagonals
//              (framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:45-161),
//              loading d*tr(S)/L (tr = sum of real parts)
//   solve      Cholesky U^H U = S_L, y = S_L^-1 a with a = 1
//              (framework/beamformer/capon/getCaponC.pyx:259-289; LA/linearSolvec.pyx:259-346)
//   outputs    w = y / (a^H y), z = w^H g with g_l = (1/K) sum_k x_c[k+l],
//              P = 1 / (K(2Y+1) a^H S_L^-1 a)   (RealTimeCapon/capon_kernel.cu:338-427)
#include "cos_device.h"
#include "capon_kernels.h"

namespace cosk {

namespace {

constexpr int WIDE_LMAX = 64;
constexpr int WIDE_NU = WIDE_LMAX * (WIDE_LMAX + 1) / 2;

// column of packed index e = r + c(c+1)/2 (r <= c)
__device__ __forceinline__ int packed_col(int e) {
  int c = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
  while (c * (c + 1) / 2 > e) --c;
  while ((c + 1) * (c + 2) / 2 <= e) ++c;
  return c;
}

__global__ __launch_bounds__(64) void capon_wide_f64_kernel(const CaponLaunch a, const int MP) {
  extern __shared__ float2 lds[];
  const int Y = a.Y, TW = 2 * Y + 1, Nzp = a.Nz - 2 * Y;
  float2* const xs = lds;                                                     // [TW][MP]
  cplx<double>* const A = reinterpret_cast<cplx<double>*>(lds + TW * MP);    // packed upper
  double* const invd = reinterpret_cast<double*>(A + WIDE_NU);               // 1/U_kk
  const int lane = threadIdx.x;
  const int npix = a.pcount ? min(*a.pcount, a.npix) : a.npix;

  for (int gid = blockIdx.x; gid < npix; gid += gridDim.x) {
    const int p = a.plist ? a.plist[gid] : (gid / a.jn) * Nzp + a.jlo + gid % a.jn;
    const int ix = p / Nzp, j = p - ix * Nzp, crow = j + Y;
    int L = a.rowL ? a.rowL[crow] : a.L;
    int M = a.rowM ? a.rowM[crow] : a.M;
    int off = a.rowOff ? a.rowOff[crow] : a.off;
    float dload = a.rowD ? a.rowD[crow] : a.d;
    // sizes that break the launch's promises: a harmless 1 x 1 problem, NaN outputs
    const bool valid = L >= 1 && L <= WIDE_LMAX && L <= a.Lcap && M >= L && M <= a.Mcap && M <= MP &&
                       off >= 0 && off + M <= a.Mstride && dload >= 0.f;
    if (!valid) {
      L = 1;
      M = 1;
      off = 0;
      dload = 1.f;
    }
    const int K = M - L + 1;

    __syncthreads();   // the previous pixel's reads of LDS are complete
    for (int t = 0; t < TW; ++t) {
      const float2* src = a.x + ((size_t)ix * a.Nz + (j + t)) * a.Mstride + off;
      for (int m = lane; m < MP; m += 64) xs[t * MP + m] = (m < M) ? src[m] : make_float2(0.f, 0.f);
    }
    __syncthreads();

    // covariance: first row by direct sums, then the displacement recursion
    // S[r][c] = S[r-1][c-1] + sum_t (x_t[K-1+r] conj(x_t[K-1+c]) - x_t[r-1] conj(x_t[c-1])),
    // lane c owning column c (rows r <= c); exact in FP64 to rounding
    const int c = lane;
    const int ucol = c * (c + 1) / 2;
    if (c < L) {
      cplx<double> acc = {0.0, 0.0};
      for (int t = 0; t < TW; ++t) {
        const float2* xr = xs + t * MP;
        for (int k = 0; k < K; ++k) {
          const float2 u = xr[k], v = xr[k + c];
          add_mul_conj(acc, cplx<double>{u.x, u.y}, cplx<double>{v.x, v.y});
        }
      }
      A[ucol] = acc;
    }
    for (int r = 1; r < L; ++r) {
      __syncthreads();   // row r-1 is complete
      if (c >= r && c < L) {
        cplx<double> acc = A[(r - 1) + (c - 1) * c / 2];
        for (int t = 0; t < TW; ++t) {
          const float2* xr = xs + t * MP;
          const float2 u1 = xr[K - 1 + r], a1 = xr[K - 1 + c], u0 = xr[r - 1], b0 = xr[c - 1];
          add_mul_conj(acc, cplx<double>{u1.x, u1.y}, cplx<double>{a1.x, a1.y});
          sub_mul_conj(acc, cplx<double>{u0.x, u0.y}, cplx<double>{b0.x, b0.y});
        }
        A[r + ucol] = acc;
      }
    }
    __syncthreads();

    // forward-backward averaging (CAP/getCaponAllC.pyx:357-358): entry (r, c) and its
    // partner (L-1-c, L-1-r) both become their mean; the pair is handled by the lane of
    // the smaller column (ties: smaller row), so no entry is written twice
    if (a.fb) {
      if (c < L) {
        for (int r = 0; r <= c; ++r) {
          const int rp = L - 1 - c, cp = L - 1 - r;
          if (c < cp || (c == cp && r < rp)) {
            const int e1 = r + ucol, e2 = rp + cp * (cp + 1) / 2;
            const cplx<double> u = A[e1], v = A[e2];
            const cplx<double> m = {0.5 * (u.x + v.x), 0.5 * (u.y + v.y)};
            A[e1] = m;
            A[e2] = m;
          }
        }
      }
      __syncthreads();
    }

    double tp = 0.0;
    if (lane < L) tp = A[lane + lane * (lane + 1) / 2].x;
    const double trS = Grp<64>::sum(tp);
    const double delta = (double)dload * trS / (double)L;

    // optional materialised R in the reference layout: mem[r*L+c] = conj(S_L[r][c])
    if (a.R && valid) {
      float2* Rp = a.R + (size_t)p * a.Rstride;
      const int NE = L * (L + 1) / 2;
      for (int e = lane; e < NE; e += 64) {
        const int cc = packed_col(e), rr = e - cc * (cc + 1) / 2;
        cplx<double> v = A[e];
        if (rr == cc) v.x += delta;
        Rp[rr * L + cc] = make_float2((float)v.x, (float)-v.y);
        if (rr < cc) Rp[cc * L + rr] = make_float2((float)v.x, (float)v.y);
      }
    }
    if (a.cov_only) continue;

    // Cholesky U^H U = S + delta I, right-looking; lane c owns column c
    bool notpd = false;
    for (int k = 0; k < L; ++k) {
      double piv = A[k + k * (k + 1) / 2].x + delta;
      if (!(piv > 1.0e-300)) {
        notpd = true;
        piv = 1.0e-300;
      }
      const double inv = 1.0 / sqrt(piv);
      cplx<double> ukc = {0.0, 0.0};
      if (lane > k && lane < L) {
        ukc = A[k + ucol];
        ukc.x *= inv;
        ukc.y *= inv;
        A[k + ucol] = ukc;
      }
      if (lane == 0) invd[k] = inv;
      __syncthreads();
      for (int i = k + 1; i < L; ++i) {
        const cplx<double> u = A[k + i * (i + 1) / 2];   // U[k][i], same for every lane
        if (lane >= i && lane < L) {
          cplx<double> v = A[i + ucol];
          sub_conj_mul(v, u, ukc);
          A[i + ucol] = v;
        }
      }
      __syncthreads();
    }

    // forward solve U^H v = a (lane c keeps v_c), |v|^2 = a^H S_L^-1 a
    cplx<double> b = {lane < L ? 1.0 : 0.0, 0.0}, vv = {0.0, 0.0};
    double accP = 0.0;
    for (int k = 0; k < L; ++k) {
      const double inv = invd[k];
      const cplx<double> vk = {rdlane(b.x, k) * inv, rdlane(b.y, k) * inv};
      accP = fma(vk.x, vk.x, fma(vk.y, vk.y, accP));
      if (lane == k) vv = vk;
      if (lane > k && lane < L) sub_conj_mul(b, A[k + ucol], vk);
    }
    const double norm = (double)K * (double)TW;
    const float nanv = __builtin_nanf("");
    if (lane == 0 && a.P && a.pmode == 0) a.P[p] = valid ? (float)(1.0 / (norm * accP)) : nanv;

    if (a.w || a.z || (a.pmode == 1 && a.P)) {
      // back solve U y = v, column-oriented
      cplx<double> cr = vv, y = {0.0, 0.0};
      for (int jj = L - 1; jj >= 0; --jj) {
        const double inv = invd[jj];
        const cplx<double> yj = {rdlane(cr.x, jj) * inv, rdlane(cr.y, jj) * inv};
        if (lane == jj) y = yj;
        if (lane < jj) {
          const cplx<double> u = A[lane + jj * (jj + 1) / 2];   // U[lane][jj]
          cplx<double> t = cr;
          add_mul(t, cplx<double>{-u.x, -u.y}, yj);
          cr = t;
        }
      }
      const double sc = 1.0 / accP;
      const cplx<double> wv = {y.x * sc, y.y * sc};
      cplx<double> zp = {0.0, 0.0};
      if (lane < L) {
        cplx<double> g = {0.0, 0.0};
        const float2* xc = xs + Y * MP + lane;
        for (int k = 0; k < K; ++k) {
          g.x += xc[k].x;
          g.y += xc[k].y;
        }
        g.x /= (double)K;
        g.y /= (double)K;
        add_conj_mul(zp, wv, g);
        if (a.w) a.w[(size_t)p * a.wstride + lane] = make_float2((float)wv.x, (float)wv.y);
      }
      zp.x = Grp<64>::sum(zp.x);
      zp.y = Grp<64>::sum(zp.y);
      if (a.pmode == 1 || a.zmode == 1) {
        const cplx<double> wl = lane < L ? wv : cplx<double>{0.0, 0.0};
        const double wn = Grp<64>::sum(fma(wl.x, wl.x, wl.y * wl.y));
        // |w^H R_hat w| = |w^H S_L w - delta |w|^2| / norm, w^H S_L w = 1/(a^H S_L^-1 a)
        if (a.pmode == 1 && lane == 0 && a.P)
          a.P[p] = valid ? (float)(fabs(1.0 / accP - delta * wn) / norm) : nanv;
        if (a.zmode == 1) {
          // power Capon: sqrt(sum_k |w^H u_{c,k}|^2) / K (RTC/capon_kernel.cu:235-335)
          __syncthreads();
          cplx<double>* const wb = A;   // the matrix is no longer needed
          wb[lane] = wl;
          __syncthreads();
          double pk = 0.0;
          const float2* xc = xs + Y * MP;
          for (int k = lane; k < K; k += 64) {
            cplx<double> sk = {0.0, 0.0};
            for (int l = 0; l < L; ++l) add_conj_mul(sk, wb[l], cplx<double>{xc[k + l].x, xc[k + l].y});
            pk = fma(sk.x, sk.x, fma(sk.y, sk.y, pk));
          }
          pk = Grp<64>::sum(pk);
          zp = {sqrt(pk) / (double)K, 0.0};
        }
      }
      if (lane == 0 && a.z) a.z[p] = valid ? make_float2((float)zp.x, (float)zp.y) : make_float2(nanv, nanv);
    }
    if (a.flags && lane == 0) {
      uint8_t f = 0;
      if ((1.0 + (double)L / (double)dload) > (double)a.cond_thresh) f |= 1;
      if (notpd) f |= 2;
      if (!valid) f = 4;
      a.flags[p] = f;
    }
  }
}

}  // namespace

hipError_t launch_capon_wide_f64(const CaponLaunch& a, int Mmax, hipStream_t s) {
  if (!a.x) return hipErrorInvalidValue;
  const int MP = (Mmax + 1) & ~1;   // even, so the matrix after the rows is 16-byte aligned
  const int TW = 2 * a.Y + 1;
  const size_t lds = (size_t)TW * MP * sizeof(float2) + WIDE_NU * sizeof(cplx<double>) + WIDE_LMAX * sizeof(double);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  int blocks = a.npix < 8192 ? a.npix : 8192;   // grid-stride over pixels
  if (blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(capon_wide_f64_kernel, dim3(blocks), dim3(64), lds, s, a, MP);
  return hipGetLastError();
}

}  // namespace cosk
