/*
 * cos_capon.h -- C ABI of the MI355X Capon/MVDR beamformer library (libcos_capon.so).
 *
 * Plain pointers and sizes only.  Complex numbers are {float re, im} interleaved,
 * layout-identical to the reference's Complex<float> (framework/lib/Complex.h:11-164),
 * hipFloatComplex and std::complex<float>.
 *
 * Data layout of every cube is the reference's x[Nx][Nz][M] (M fastest,
 * framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:18,86).  Pixel
 * p = ix*(Nz-2*Yavg) + j; edge rows are dropped, not padded
 * (RealTimeCapon/Capon.cpp:90-91).
 *
 * Return codes: 0 = success; otherwise a HIP error number (1 = invalid argument,
 * 2 = out of device memory, others from the HIP runtime).  cos_capon_last_error()
 * returns a message for the last failure on the calling thread.  No function
 * writes files or prints.
 */
#ifndef COS_CAPON_H
#define COS_CAPON_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float re, im; } cos_cf32;

/* ----------------------------------------------------------------------------------
 * Drop-in entry points (host pointers; the library copies to and from the GPU).
 * ---------------------------------------------------------------------------------- */

/* Replaces Capon::getCapon (RealTimeCapon/ICapon.h:20-33, RealTimeCapon/Capon.cpp:73-152).
 * z [Nx][Nz-2Y] amplitude, w [Nx][Nz-2Y][L] weights (element space also when Nb > 0),
 * R [Nx][Nz-2Y][L][L] the loaded, unnormalised covariance stored conjugated (= column
 * major), as the reference copies back with COPY_R_DEVICE_HOST (Capon.cpp:9,255-260);
 * R may be NULL to skip it.  d is the relative diagonal loading (d*trace/L).
 * Nb > 0 selects the reference's default matrix beamspace path with Nb Butler beams
 * (Capon.cpp:285-402).  Requirements as the reference: 1 <= L <= M, Yavg >= 0,
 * 2*Yavg < Nz, 0 <= Nb <= L, L <= 72 (the solver's limit, Solver.cpp:34-42); for
 * 64 < L <= 72 (Nb = 0) a kernel with one block per pixel in LDS serves the call. */
int cos_capon_get_capon(cos_cf32* z, cos_cf32* w, cos_cf32* R, const cos_cf32* x, float d,
                        int L, int Yavg, int M, int Nx, int Nz, int Nb);

/* As cos_capon_get_capon, plus the power image P [Nx][Nz-2Y] = 1/(a^H R_L^-1 a) with
 * R_L = S_L/(K(2Y+1)); no reference counterpart (the intended zPow of
 * framework/beamformer/capon/getCaponC.pyx:285-289).  Any output may be NULL. */
int cos_capon_get_capon_p(cos_cf32* z, cos_cf32* w, cos_cf32* R, float* P, const cos_cf32* x,
                          float d, int L, int Yavg, int M, int Nx, int Nz, int Nb);

/* Replaces BuildR::getR (RealTimeCapon/IBuildR.h:51-64, RealTimeCapon/BuildR.cpp:21-106).
 * Nb == 0: R [Nx][Nz-2Y][L][L] loaded unnormalised conj-stored covariance of x [Nx][Nz][M].
 * Nb > 0: x holds beamspace-transformed subarrays [Nx][Nz][K][Nb] (K = M-L+1) and R is
 * [Nx][Nz-2Y][Nb][Nb] with loading d*trace/Nb (BuildR.cpp:82-86, buildR_kernel.cu:169-251).
 * R_on_gpu / x_on_gpu: nonzero when the pointer is device memory.  L <= 128 (Nb = 0). */
int cos_buildr_get_r(cos_cf32* R, const cos_cf32* x, float d, int L, int Yavg, int M, int Nx,
                     int Nz, int Nb, int R_on_gpu, int x_on_gpu);

/* Replaces Solver::solve (RealTimeCapon/ISolver.h:16-29, RealTimeCapon/Solver.cpp:23-116):
 * batch general complex systems A x = b, A column-major N x N (as the reference's kernels
 * read it, RealTimeCapon/nvidia_solve.h:44-46), 1 <= N <= 72.  Gaussian elimination with
 * partial pivoting.  *_on_gpu: nonzero when the pointer is device memory. */
int cos_solver_solve(cos_cf32* x, const cos_cf32* A, const cos_cf32* b, int N, int batch,
                     int x_on_gpu, int A_on_gpu, int b_on_gpu);

/* Message for the last failure on this thread ("" if none). */
const char* cos_capon_last_error(void);

/* ----------------------------------------------------------------------------------
 * Plan API (device pointers, caller's stream).  No reference counterpart: the
 * reference keeps device buffers inside its Capon object (Capon.h:16-40); a plan is
 * that object made explicit, one per thread of use.
 * ---------------------------------------------------------------------------------- */

typedef struct cos_capon_plan cos_capon_plan;

enum {
    COS_PREC_FP32 = 0,           /* FP32 covariance and solve */
    COS_PREC_FP64 = 1,           /* plain per-pixel FP64: direct covariance, FP64 Cholesky */
    COS_PREC_FP32_FALLBACK = 2   /* FP32, then FP64 for pixels flagged ill-conditioned */
};

typedef struct {
    int Nx, Nz;          /* beams, range samples */
    int Mstride;         /* complex elements per snapshot row in memory */
    int M, L, off;       /* active elements, subarray length, first active element */
    int Yavg;            /* temporal window 2*Yavg+1 */
    float d;             /* relative diagonal loading */
    const int* row_L;    /* optional device tables of length Nz indexed by the centre row */
    const int* row_M;    /*   (override L, M, off, d per range row) */
    const int* row_off;
    const float* row_d;
    int Lmax, Mmax;      /* bounds on the row tables (ignored without tables) */
    int precision;       /* COS_PREC_* */
    float cond_thresh;   /* fallback when 1 + L/d exceeds this (default 1e6 when <= 0) */
    const float* steer_sin; /* optional device array of nsteer sin(theta) values */
    int nsteer;
    int graph;           /* nonzero: capture the frame's launches in a hipGraph on the first
                            call and replay it while descriptor and pointers are unchanged
                            (the row tables' contents may change: rows are sorted into lane
                            shapes on the device inside the graph) */
    /* Options of the reference's CPU path (framework/beamformer/capon/getCaponAllC.pyx)
     * and compiled-out GPU variants; 0 selects the reference GPU default. */
    int fb_avg;          /* forward-backward averaging R <- (R + J R^T J)/2 before loading
                            (getCaponAllC.pyx:357-358) */
    int z_mode;          /* COS_Z_*: amplitude output */
    int p_mode;          /* COS_P_*: power output */
    int refine;          /* nonzero: one step of iterative refinement on the FP32 paths, the
                            residual a - R_L y taken from the snapshots themselves (a = 1 or
                            the steering vector; no forward-backward averaging) */
    int cov_mfma;        /* 0: covariance by the FP32 displacement recursion, its rank-two
                            corrections on the matrix cores where 3 applies (the default);
                            1: covariance on the matrix cores (BF16 hi/lo split, FP32
                            accumulation) into R (or a plan buffer when R is NULL), then the
                            FP32 solve from it; L <= 64, row tables allowed (R then a plan
                            buffer of Lmax*Lmax per pixel), default options;
                            2: per-snapshot Grams on the matrix cores inside the fused kernel
                            (FP16 hi/lo split of power-of-two scaled rows, FP32 accumulation),
                            shared by the windows of a wave's four consecutive pixels, where
                            it applies (16 lanes x 2 columns, 16 < L <= 32, windows of up to
                            five rows, default options, no R output); the recursion elsewhere;
                            3: the displacement recursion with its rank-two corrections
                            (two per window row) summed over each window on the FP32 matrix
                            cores, shared by a wave's four consecutive pixels, where it
                            applies (16 lanes x 2 columns, 16 < L <= 32, windows of up to
                            five rows); no split, FP32 products and accumulation;
                            4: the displacement recursion entirely on the VALU */
} cos_capon_desc;

enum {
    COS_Z_AMPLITUDE = 0, /* z = w^H g, coherent sum over subarrays (RealTimeCapon/capon_kernel.cu:338-427) */
    COS_Z_POWER = 1      /* z = sqrt(sum_k |w^H u_{c,k}|^2) / K, the power-Capon amplitude of the
                            compiled-out power_capon_kernel (RealTimeCapon/capon_kernel.cu:235-335) */
};
enum {
    COS_P_LOADED = 0,    /* P = 1/(a^H R_L^-1 a) = w^H R_L w */
    COS_P_UNLOADED = 1   /* P = |w^H R_hat w| with the unloaded R_hat, the definition of
                            framework/beamformer/capon/getCaponMatlab_ref.m:109 */
};

int cos_capon_plan_create(cos_capon_plan** plan);
int cos_capon_plan_destroy(cos_capon_plan* plan);

/* One frame.  Outputs (device, any may be NULL): P [npx*S], z [npx*S], w [npx*S][wstride],
 * R [npx][L*L] (uniform L only), flags [npx] (bit0 condition bound above threshold,
 * bit1 non-positive pivot in FP32), with npx = Nx*(Nz-2*Yavg) and S = max(nsteer, 1).
 * stream is a hipStream_t (NULL = default stream).  Asynchronous: returns after
 * enqueueing; the FP64 fallback pass is enqueued on the same stream. */
int cos_capon_run(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x,
                  float* P, cos_cf32* z, cos_cf32* w, int wstride, cos_cf32* R,
                  unsigned char* flags, void* stream);

/* As cos_capon_run with no plan and no allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The FP64 fallback's scratch (flags when the
 * caller passes none, the pixel list and its count) is a fixed device array of
 * 256 + 5 * COS_CAPON_NOALLOC_MAX_PIXELS bytes per GPU, held until the process ends, so a
 * call takes at most COS_CAPON_NOALLOC_MAX_PIXELS output pixels (larger: error 1).
 * Uniform L, M and d only (no row tables), no graph capture, no matrix-core covariance,
 * no R output; every other descriptor field as cos_capon_run.  Calls that overlap on one
 * GPU share its scratch and race: serialise them (one stream, or one call at a time). */
#define COS_CAPON_NOALLOC_MAX_PIXELS 2097152
int cos_capon_run_noalloc(const cos_capon_desc* desc, const cos_cf32* x, float* P, cos_cf32* z,
                          cos_cf32* w, int wstride, unsigned char* flags, void* stream);

/* Autotuner of the fused FP32 kernel's lane shape (no reference counterpart; the
 * reference has one kernel configuration).  For each subarray length the frame uses (the
 * uniform L, or each distinct value of the row table) it times every lane shape that holds
 * that length (1: 16 lanes x 1 column, L <= 16; 2: 16 lanes x 2 columns, L <= 32;
 * 3: 64 lanes, L <= 64; 4: 16 lanes x 3 columns, L <= 48) on up to 128
 * output rows of that length (all beams) and keeps the fastest for this device and frame
 * configuration (Mmax, Yavg, options, refine); later cos_capon_run calls on the device use
 * it.  P, z, w (device, nullable, as cos_capon_run) receive the timed rows' results.
 * shapes_out [65] (nullable): chosen shape per L, 0 where the frame has no row of that
 * length; ms_out [65*COS_CAPON_SHAPES] (nullable): ms per launch of shape s at L in
 * ms_out[COS_CAPON_SHAPES*L + s - 1], negative where not timed.  Synchronous on stream. */
#define COS_CAPON_SHAPES 4
int cos_capon_autotune(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x, float* P,
                       cos_cf32* z, cos_cf32* w, int wstride, void* stream, int* shapes_out, float* ms_out);

/* Installs a lane-shape table shapes[65] for the current device and the frame
 * configuration of desc, e.g. one saved from cos_capon_autotune.  0 = untuned: the 16-lane
 * shape with the fewest columns that holds L, 64 lanes past L = 48.  Rejects a shape that
 * does not hold its L. */
int cos_capon_tuning_set(const cos_capon_desc* desc, const int* shapes);

/* Beamspace Capon with the sliding DFT: replaces the reference's SLIDING_BEAMSPACE build
 * (RealTimeCapon/sliding_beamspace_kernel.cu:26-237 -> build_R_full -> Nb-solve ->
 * getCaponOutputFromSubarraySums, RealTimeCapon/Capon.cpp:202-227,399-415) with loading
 * d*tr(R_bs)/Nb as in framework/beamformer/capon/getCaponAllC.pyx:376-385.  Uses Nx, Nz,
 * Mstride, M, L, off, Yavg, d of desc; 1 <= Nb <= 8 Butler beams symmetric about broadside.
 * Outputs (device, any may be NULL): P [npx] power 1/(a_bs^H R_bs^-1 a_bs), z [npx]
 * amplitude w_bs^H (B g), w_bs [npx][Nb], w [npx][wstride] element-space weights B^H w_bs. */
int cos_capon_beamspace_run(cos_capon_plan* plan, const cos_capon_desc* desc, int Nb,
                            const cos_cf32* x, float* P, cos_cf32* z, cos_cf32* w_bs,
                            cos_cf32* w, int wstride, void* stream);

/* Capon in the subspace spanned by the rows of a user matrix V [Nb][L] (device, row-major):
 * the reference CPU path's subspace option (framework/beamformer/capon/getCaponAllC.pyx:
 * 230,346-447, loading d*tr(R)/Nb in the subspace, its default), which its GPU build does
 * not implement (only Nb is taken from V, framework/beamformer/capon/getCaponCUDA.pyx:33-34).
 * R = V R_hat V^H with R_hat the (forward-backward averaged when desc.fb_avg) normalised
 * covariance, a = V 1, g = V g; uses Nx, Nz, Mstride, M, L, off, Yavg, d, fb_avg of desc.
 * Outputs (device, any may be NULL): P [npx] 1/(a^H R_L^-1 a), z [npx] w^H g,
 * w [npx][Nb] subspace weights.  Keeps an [npx][L*L] covariance buffer in the plan. */
int cos_capon_subspace_run(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* V, int Nb,
                           const cos_cf32* x, float* P, cos_cf32* z, cos_cf32* w, void* stream);

/* Adjoint of the power image (no reference counterpart; the differentiable workload).
 * Given dl/dP (grad_P [npx], device), adds dl/dx into grad_x (same layout as x; the caller
 * zeroes it first; accumulated with atomics) and writes the per-pixel dl/dd into grad_d
 * [npx] (may be NULL; sum it for a frame-wide d, or per row for row tables).  The
 * covariance and its Cholesky factor are recomputed, not stored.  Gradients follow the
 * d/dRe + j d/dIm convention.  FP32, L <= 32, M <= 128, no steering. */
int cos_capon_backward(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x,
                       const float* grad_P, cos_cf32* grad_x, float* grad_d, void* stream);

/* The stages either side of the beamformer call.
 *
 * Delay stage: time-domain interpolation beamformer, replacing
 * framework/beamformer/delay/tdibC.pyx:9-67 (tdib(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr)).
 * mfdata [n_h][n_t] matched-filtered channels on the uniform axis t0 + i*dt (any strictly
 * increasing axis: cos_tdib_run_axis); x_r, D [n_h] hydrophone positions;
 * xarr, yarr [n_x][n_y] pixel coordinates; all device pointers.  Writes the delayed,
 * carrier-mixed channels cube [n_x][n_y][n_h] (the Capon input layout) and/or the
 * delay-and-sum image [n_y][n_x] = (1/n_h) sum over hydrophones (the reference's return
 * value); either may be NULL.  Times outside the axis give `fill` (the reference: NaN). */
int cos_tdib_run(const cos_cf32* mfdata, int n_h, int n_t, double t0, double dt, const double* x_r,
                 const double* D, double c, double fc, const double* xarr, const double* yarr,
                 int n_x, int n_y, float fill, cos_cf32* cube, cos_cf32* image, void* stream);

/* As cos_tdib_run on any strictly increasing time axis taxis [n_t] (device, seconds), as
 * the reference's interp1d takes (tdibC.pyx:55-56): the interval of t is scipy's, the first
 * sample at or after t (searchsorted, left side) clipped to 1 .. n_t-1 and the one before it,
 * found by binary search; times outside [taxis[0], taxis[n_t-1]] give `fill`. */
int cos_tdib_run_axis(const cos_cf32* mfdata, int n_h, int n_t, const double* taxis, const double* x_r,
                      const double* D, double c, double fc, const double* xarr, const double* yarr,
                      int n_x, int n_y, float fill, cos_cf32* cube, cos_cf32* image, void* stream);

/* Window selection: low-complexity adaptive beamformer, replacing
 * framework/beamformer/lca/lcacx.pyx:27-217 (lca(Z, wrf)).  Z [Nx][Ny][Nw] images formed
 * with Nw windows; wrf [wy][wx] neighbourhood weights (odd sizes), device pointers.
 * img [Nx][Ny] gets Z at the selected window (the first minimum of the weighted
 * neighbourhood power), win [Nx][Ny] its index; either output may be NULL.
 * mode COS_LCA_REFERENCE: the reference's output index for index, in its own branches
 *   (chosen from the filter's values, which are read to the host: the call synchronises
 *   the stream): a 1 x 1 filter with a nonzero weight selects argmin |Z| at every pixel;
 *   an all-ones range filter (wx = 1) selects by FP64 sliding sums at rows
 *   Ym+1 .. Ny-Ym-1, writes window 0 at row wy when that row is past them, 0 elsewhere
 *   (row Ym included); any other filter reads the |Z|^2 array with x and y swapped
 *   (lcacx.pyx:206; the transposed neighbourhood on square images) and writes the
 *   interior.  Shapes where the reference reads or writes outside its arrays are refused:
 *   Ny <= wy for the all-ones filter, Ny > Nx for the others.  Nw <= 256 for the all-ones
 *   filter.
 * mode COS_LCA_CENTRED: the intended neighbourhood sum_{yy,xx} wrf[yy][xx]
 *   |Z[x+xx-Xm][y+yy-Ym][w]|^2 (FP32) at every pixel whose neighbourhood lies inside the
 *   image, 0 and window 0 elsewhere; any shape. */
enum { COS_LCA_REFERENCE = 0, COS_LCA_CENTRED = 1 };
int cos_lca_run(const cos_cf32* Z, int Nx, int Ny, int Nw, const float* wrf, int wy, int wx, int mode,
                cos_cf32* img, int* win, void* stream);

/* Number of pixels the last cos_capon_run sent to the FP64 fallback (synchronises the
 * plan's stream). */
int cos_capon_last_fallback_count(cos_capon_plan* plan, int* count);

#ifdef __cplusplus
}
#endif
#endif /* COS_CAPON_H */
