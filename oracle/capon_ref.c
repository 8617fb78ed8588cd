/*
 * CPU FP64 restatement of the reference's per-pixel Capon loop.
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/liboracle_capon.so and loaded by
 * tests/ (as a second checker) and by bench.py's cpu_baseline leg, which times
 * it on the host's cores beside the GPU.  Nothing in the product links it.
 *
 * It follows framework/beamformer/capon/getCaponC.pyx:196-289 step by step:
 *   gather the 2Y+1 snapshots (:203-205), banded R_full[i][i+j] =
 *   sum_t conj(ar[t][i]) ar[t][i+j] (:208-223), subarray sum of the upper triangle
 *   (:227-237), conj and 1/(K(2Y+1)) (:240-242), g by sliding subarray sums
 *   (:247-255), loading tr*d/L (:258-262), Cholesky U^H U
 *   (framework/linalg/matrixDecompc.pyx:19-50), forward/back solves
 *   (framework/linalg/linearSolvec.pyx:312-346), w = y/(a^H y) and z = w^H g
 *   (:276-283).  P = 1/Re(a^H y) is added (the intended zPow).
 * Arithmetic is double complex throughout; d, 1/K and 1/norm are exact doubles.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cd;

static void capon_pixel(const float *x, int Nz, int M, int L, int Y, double d,
                        int ix, int j, cd *ar, cd *Rfull, cd *R, cd *U, cd *tmp,
                        double *P, double *zo, double *wo)
{
    const int T = 2 * Y + 1, K = M - L + 1;
    const double norm_inv = 1.0 / ((double)K * T), K_inv = 1.0 / K;
    for (int t = 0; t < T; ++t)
        for (int m = 0; m < M; ++m) {
            const float *p = x + 2 * (((size_t)ix * Nz + (j + t)) * M + m);
            ar[t * M + m] = p[0] + I * p[1];
        }
    /* banded R_full[i][i+jj] for jj < L (getCaponC.pyx:208-223) */
    for (int i = 0; i < M; ++i)
        for (int jj = 0; jj < L && i + jj < M; ++jj) {
            cd s = 0;
            for (int t = 0; t < T; ++t) s += conj(ar[t * M + i]) * ar[t * M + i + jj];
            Rfull[i * L + jj] = s;
        }
    /* subarray sum, upper triangle (getCaponC.pyx:227-237), conj + normalise (:240-242) */
    for (int a = 0; a < L; ++a)
        for (int b = a; b < L; ++b) {
            cd s = 0;
            for (int i = 0; i < K; ++i) s += Rfull[(i + a) * L + (b - a)];
            R[a * L + b] = conj(s) * norm_inv;
        }
    /* g: subarray sums of the centre snapshot (getCaponC.pyx:247-255) */
    const cd *xc = ar + Y * M;
    cd s0 = 0;
    for (int i = 0; i < K; ++i) s0 += xc[i];
    cd *g = tmp + 2 * L;
    g[0] = s0 * K_inv;
    for (int i = 1; i < L; ++i) {
        s0 = s0 - xc[i - 1] + xc[M - L + i];
        g[i] = s0 * K_inv;
    }
    /* loading (getCaponC.pyx:258-262) */
    cd tr = 0;
    for (int i = 0; i < L; ++i) tr += R[i * L + i];
    const cd dl = tr * d / L;
    for (int i = 0; i < L; ++i) R[i * L + i] += dl;
    /* Cholesky U^H U = R (matrixDecompc.pyx:19-50) */
    memset(U, 0, sizeof(cd) * L * L);
    for (int i = 0; i < L; ++i) {
        cd acc = 0;
        for (int k = 0; k < i; ++k) acc += U[k * L + i] * conj(U[k * L + i]);
        U[i * L + i] = csqrt(R[i * L + i] - acc);
        for (int c = i + 1; c < L; ++c) {
            acc = 0;
            for (int k = 0; k < i; ++k) acc += conj(U[k * L + i]) * U[k * L + c];
            U[i * L + c] = (R[i * L + c] - acc) / U[i * L + i];
        }
    }
    /* forward U^H v = a (a = 1), back U y = v (linearSolvec.pyx:312-346) */
    cd *v = tmp, *y = tmp + L;
    for (int i = 0; i < L; ++i) {
        cd acc = 1.0;
        for (int k = 0; k < i; ++k) acc -= conj(U[k * L + i]) * v[k];
        v[i] = acc / conj(U[i * L + i]);
    }
    for (int i = L - 1; i >= 0; --i) {
        cd acc = v[i];
        for (int k = L - 1; k > i; --k) acc -= U[i * L + k] * y[k];
        y[i] = acc / U[i * L + i];
    }
    cd aRia = 0;
    for (int i = 0; i < L; ++i) aRia += y[i];
    cd z = 0;
    for (int i = 0; i < L; ++i) {
        cd w = y[i] / aRia;
        if (wo) { wo[2 * i] = creal(w); wo[2 * i + 1] = cimag(w); }
        z += conj(w) * g[i];
    }
    *P = 1.0 / creal(aRia);
    zo[0] = creal(z);
    zo[1] = cimag(z);
}

/* Pixels (ix, j) for ix in [ix0, ix1), j in [j0, j1); outputs are [ix-ix0][j-j0].
 * w may be NULL.  Returns 0, or -1 on bad arguments. */
int oracle_capon_f64(const float *x, int Nx, int Nz, int M, int L, int Yavg, double d,
                     int ix0, int ix1, int j0, int j1,
                     double *P, double *z, double *w, int nthreads)
{
    if (L < 1 || L > M || Yavg < 0 || 2 * Yavg >= Nz || ix0 < 0 || ix1 > Nx || ix0 > ix1 ||
        j0 < 0 || j1 > Nz - 2 * Yavg || j0 > j1)
        return -1;
    const int nj = j1 - j0, npx = (ix1 - ix0) * nj, T = 2 * Yavg + 1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        cd *ar = malloc(sizeof(cd) * T * M);
        cd *Rfull = malloc(sizeof(cd) * M * L);
        cd *R = malloc(sizeof(cd) * L * L);
        cd *U = malloc(sizeof(cd) * L * L);
        cd *tmp = malloc(sizeof(cd) * 3 * L);
#pragma omp for schedule(dynamic, 16)
        for (int p = 0; p < npx; ++p) {
            int ix = ix0 + p / nj, j = j0 + p % nj;
            capon_pixel(x, Nz, M, L, Yavg, d, ix, j, ar, Rfull, R, U, tmp, P + p, z + 2 * p,
                        w ? w + 2 * (size_t)p * L : NULL);
        }
        free(ar); free(Rfull); free(R); free(U); free(tmp);
    }
    return 0;
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
