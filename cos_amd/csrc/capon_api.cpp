// Host side of the C ABI declared in include/cos_capon.h.
//
// Drop-in entry points take host pointers, stage them through device buffers owned
// by a process-wide plan (grown on demand, never shrunk; the reference re-allocates
// on every call because cur_Npx is never updated, RealTimeCapon/Capon.cpp:110-141,
// :528-547), run on the null stream and copy the outputs back.  Plan entry points
// take device pointers and a stream and only enqueue work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cos_capon.h"
#include "capon_kernels.h"
#include "beamspace.h"
#include "components.h"
#include "aux_stages.h"

static_assert(cosk::kShapes == COS_CAPON_SHAPES, "the ABI's lane-shape count");

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
// The error is reported here and cleared from HIP's per-thread last error, which a
// failed runtime call (an allocation past the device's memory) leaves set: the next
// launch's hipGetLastError would otherwise report it again.
int fail_hip(hipError_t e, const char* where) {
  g_last_error = std::string(where) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return (int)e;
}

// Grow-only device buffer.  A reallocation bumps *gen (the owning plan's buffer
// generation), so a frame graph that captured the old address is never replayed.
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  unsigned* gen = nullptr;
  hipError_t ensure(size_t bytes) {
    if (bytes <= n) return hipSuccess;
    if (gen) ++*gen;
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipSuccess) n = bytes;
    else p = nullptr;
    return e;
  }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

// Autotuned lane shapes: for each device and each frame configuration that changes the
// kernels' cost (largest aperture, window, options), the fastest shape per subarray
// length L = 1..64 as measured by cos_capon_autotune (0 = not tuned: default_shape,
// the 16-lane shape with the fewest columns that holds L, 64 lanes past L = 48).  Process-wide, so every plan on the device uses it; g_tune_gen counts
// changes so captured frame graphs are re-captured after one.
using TuneKey = std::tuple<int, int, int, int, int, int, int>;   // device, Mmax, Yavg, fb, z, p, refine
using TuneTable = std::array<unsigned char, 65>;
std::mutex g_tune_mu;
std::map<TuneKey, TuneTable> g_tune;
std::atomic<int> g_tune_gen{0};

TuneKey tune_key(const cos_capon_desc& ds) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  return TuneKey{dev, tables ? ds.Mmax : ds.M, ds.Yavg, ds.fb_avg ? 1 : 0, ds.z_mode, ds.p_mode, ds.refine ? 1 : 0};
}

// Shapes for this frame: the tuned table of its key, or all zeros.
TuneTable tuned_shapes(const cos_capon_desc& ds) {
  TuneTable t{};
  if (ds.precision == COS_PREC_FP64) return t;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(tune_key(ds));
  if (it != g_tune.end()) t = it->second;
  return t;
}

inline int shape_for(const TuneTable& t, int L) {
  if (L < 1 || L > 64) return cosk::default_shape(L < 1 ? 1 : 64);
  const int s = t[L];
  return s >= 1 && s <= cosk::kShapes && cosk::shape_fits(s, L) ? s : cosk::default_shape(L);
}

// The frame's launches captured once and replayed: valid while the descriptor, every
// pointer, wstride, the tuning table and the plan's own buffers are unchanged.
struct FrameGraph {
  bool valid = false;
  cos_capon_desc ds;
  const void* ptr[7];
  int wstride = 0;
  int tune_gen = 0;
  unsigned buf_gen = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
};

struct cos_capon_plan {
  DevBuf flags, list, count;  // fallback bookkeeping
  DevBuf x, z, w, R, P, B;    // drop-in staging
  DevBuf bs_tw;               // beamspace twiddles [Nb][L] + [Nb]
  int bs_Nb = 0, bs_L = 0;
  hipStream_t last_stream = nullptr;
  hipStream_t capture_stream = nullptr;  // graphs are captured here, launched on the caller's
  FrameGraph fg;
  DevBuf rows;                // table-driven frames: per-bucket counts [8], then row lists [nb][Nzp]
  bool fallback_ran = false;
  int fb_lists = 1;     // FP64 fallback lists of the last frame (one per lane shape with row tables)
  unsigned buf_gen = 0;       // bumped by every reallocation of a plan buffer
  cos_capon_plan() {
    for (DevBuf* b : {&flags, &list, &count, &x, &z, &w, &R, &P, &B, &bs_tw, &rows}) b->gen = &buf_gen;
  }
  ~cos_capon_plan() {
    fg.reset();
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    for (DevBuf* b : {&flags, &list, &count, &x, &z, &w, &R, &P, &B, &bs_tw, &rows}) b->release();
  }
};

namespace {

std::mutex g_default_mu;
cos_capon_plan* g_default = nullptr;

cos_capon_plan* default_plan() {
  if (!g_default) g_default = new cos_capon_plan();
  return g_default;
}

// Argument checks in the reference's order (RealTimeCapon/Capon.cpp:98-107,
// RealTimeCapon/BuildR.cpp:37-48) plus this library's kernel limits.
int check_args(float d, int L, int Yavg, int M, int Nx, int Nz, int Nb) {
  if (Nx < 1 || Nz < 1 || L < 1 || M < 1 || Yavg < 0 || Nb < 0)
    return fail(hipErrorInvalidValue, "Nx or Nz were < 1, or L or M < 1, or Yavg or Nb < 0");
  if (Nb > L) return fail(hipErrorInvalidValue, "Nb can not be larger than L");
  if (L > M) return fail(hipErrorInvalidValue, "L > M");
  if (2 * Yavg >= Nz) return fail(hipErrorInvalidValue, "2*Yavg >= Nz");
  if (L > 72) return fail(hipErrorInvalidValue, "L > 72 is not supported (the reference solver's limit)");
  if (!(d >= 0.0f) || !std::isfinite(d)) return fail(hipErrorInvalidValue, "d must be finite and >= 0");
  if ((long)Nx * Nz * M > (1L << 31) - 1) return fail(hipErrorInvalidValue, "cube too large for one call");
  return 0;
}

cosk::CaponLaunch base_launch(const cos_capon_desc& ds, const void* x) {
  cosk::CaponLaunch a;
  std::memset(&a, 0, sizeof(a));
  a.x = static_cast<const float2*>(x);
  a.Nx = ds.Nx;
  a.Nz = ds.Nz;
  a.Mstride = ds.Mstride;
  a.Y = ds.Yavg;
  a.L = ds.L;
  a.M = ds.M;
  a.off = ds.off;
  a.d = ds.d;
  a.rowL = ds.row_L;
  a.rowM = ds.row_M;
  a.rowOff = ds.row_off;
  a.rowD = ds.row_d;
  a.npix = ds.Nx * (ds.Nz - 2 * ds.Yavg);
  a.jlo = 0;
  a.jn = ds.Nz - 2 * ds.Yavg;
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  a.Lcap = tables ? ds.Lmax : ds.L;
  a.Mcap = tables ? ds.Mmax : ds.M;
  a.cond_thresh = ds.cond_thresh > 0 ? ds.cond_thresh : 1.0e6f;
  a.steer_sin = ds.steer_sin;
  a.nsteer = ds.nsteer;
  // a frame with fewer pixel groups than the GPU holds waves splits its angles over the
  // grid's second dimension (each block factorises again: one factorisation costs about
  // what one refined angle does at these sizes); ~4,096 blocks in all
  if (ds.nsteer > 1) {
    const long groups = ((long)a.npix + 3) / 4;
    const int want = (int)std::min<long>(ds.nsteer, (4096 + groups - 1) / groups);
    if (want > 1) {
      const int chunk = (ds.nsteer + want - 1) / want;
      a.steer_chunk = chunk;
    }
  }
  a.fb = ds.fb_avg ? 1 : 0;
  a.zmode = ds.z_mode;
  a.pmode = ds.p_mode;
  return a;
}

// Butler beams (RealTimeCapon/capon_kernel.cu:47-74) in double on the host:
// B[b][l] = L^-1/2 exp(-j 2 pi beta(b) l / L); beta(b) = b for b <= Nb/2 else
// b + L - 2(Nb/2) - 1 (the even-Nb +1 at capon_kernel.cu:58 never fires).
std::vector<float> butler_host(int Nb, int L) {
  std::vector<float> B(2 * (size_t)Nb * L);
  const double s = 1.0 / std::sqrt((double)L);
  for (int b = 0; b < Nb; ++b) {
    int beta = b > Nb / 2 ? b + L - 2 * (Nb / 2) - 1 : b;
    for (int l = 0; l < L; ++l) {
      double ph = -2.0 * M_PI * (double)((long)beta * l % L) / (double)L;
      B[2 * ((size_t)b * L + l)] = (float)(s * std::cos(ph));
      B[2 * ((size_t)b * L + l) + 1] = (float)(s * std::sin(ph));
    }
  }
  return B;
}

}  // namespace

extern "C" {

const char* cos_capon_last_error(void) { return g_last_error.c_str(); }

int cos_capon_plan_create(cos_capon_plan** plan) {
  if (!plan) return fail(hipErrorInvalidValue, "plan is NULL");
  *plan = new cos_capon_plan();
  return 0;
}

int cos_capon_plan_destroy(cos_capon_plan* plan) {
  delete plan;
  return 0;
}

}  // extern "C"

namespace {

// Launch buckets of a table-driven frame: every subarray length 1 .. Lmax that a row may
// hold is given a bucket (its FP32 lane shape, or its FP64 kernel), one launch per bucket.
// Which rows fall in which bucket is decided on the device at every run of the frame
// (cosk::plan_rows, captured in the frame graph), so the host never reads the tables and
// a replay follows their current contents.
struct Buckets {
  cosk::RowBuckets bk{};
  int nb = 0;
  int cap[8] = {};     // largest L of the bucket's rows (its launch's Lcap)
  int shape[8] = {};   // FP32: lane shape
};

Buckets make_buckets(const cos_capon_desc& ds, const TuneTable& shapes) {
  Buckets r;
  int of[65];
  for (int L = 1; L <= ds.Lmax; ++L) {
    const int key = ds.precision == COS_PREC_FP64 ? (L <= 16 ? 1 : L <= 32 ? 2 : 3) : shape_for(shapes, L);
    int b = 0;
    while (b < r.nb && r.shape[b] != key) ++b;
    if (b == r.nb) {
      r.shape[b] = key;
      ++r.nb;
    }
    r.cap[b] = L;
    of[L] = b;
  }
  // a length outside 1 .. Lmax goes to the bucket of Lmax, whose launch writes that row as
  // NaN with flag 4 (the kernels' validity test)
  for (int L = 0; L <= 64; ++L) r.bk.b[L] = (unsigned char)(L >= 1 && L <= ds.Lmax ? of[L] : of[ds.Lmax]);
  return r;
}

// What one frame's launches need beyond the descriptor and the caller's buffers: the
// buckets and device row lists of a table-driven frame, the tuned shapes, the count of
// pixels the condition bound flags (known to the host for uniform frames), and device
// scratch (per-pixel flags when the caller passes none, the compacted FP64 pixel list and
// its device-side length, the matrix-core covariance when the caller passes no R).
struct FrameCtx {
  Buckets bu;
  int* rows = nullptr;   // [8] counts, then [nb][Nzp] row lists
  TuneTable shapes{};
  long fb_bound = 0;
  uint8_t* flags = nullptr;
  int* list = nullptr;
  int* count = nullptr;
  float2* R = nullptr;
};

// Enqueue one frame on stream s (arguments already validated; scratch already sized).
int fb_lists(const cos_capon_desc& ds);
int enqueue_frame(const FrameCtx& c, const cos_capon_desc& ds, const cos_cf32* x, float* P, cos_cf32* z,
                  cos_cf32* w, int wstride, cos_cf32* R, unsigned char* flags, hipStream_t s, bool fb_needed);

int validate(const cos_capon_desc& ds, cos_cf32* R, cos_cf32* w, int wstride) {
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  if (tables && !(ds.row_L && ds.row_M && ds.row_off && ds.row_d))
    return fail(hipErrorInvalidValue, "row tables must be given together (L, M, off, d)");
  const int Lmax = tables ? ds.Lmax : ds.L;
  const int Mmax = tables ? ds.Mmax : ds.M;
  if (int e = check_args(tables ? 0.f : ds.d, tables ? 1 : ds.L, ds.Yavg, tables ? 1 : ds.M, ds.Nx,
                         ds.Nz, 0))
    return e;
  if (Lmax < 1 || Lmax > 64 || Mmax < Lmax) return fail(hipErrorInvalidValue, "need 1 <= Lmax <= 64, Mmax >= Lmax");
  if (!tables && ds.off + ds.M > ds.Mstride) return fail(hipErrorInvalidValue, "off + M > Mstride");
  if (tables && Mmax > ds.Mstride) return fail(hipErrorInvalidValue, "Mmax > Mstride");
  if (R && tables) return fail(hipErrorInvalidValue, "R output needs a uniform L");
  if (w && wstride < Lmax) return fail(hipErrorInvalidValue, "wstride < L");
  if (ds.nsteer < 0 || (ds.nsteer > 0 && !ds.steer_sin)) return fail(hipErrorInvalidValue, "bad steering");
  if (ds.z_mode != COS_Z_AMPLITUDE && ds.z_mode != COS_Z_POWER) return fail(hipErrorInvalidValue, "bad z_mode");
  if (ds.p_mode != COS_P_LOADED && ds.p_mode != COS_P_UNLOADED) return fail(hipErrorInvalidValue, "bad p_mode");
  if (ds.nsteer > 0 && (ds.z_mode != COS_Z_AMPLITUDE || ds.p_mode != COS_P_LOADED))
    return fail(hipErrorInvalidValue, "steering takes the default z and P definitions");
  if (ds.refine && (ds.precision == COS_PREC_FP64 || ds.fb_avg))
    return fail(hipErrorInvalidValue, "refinement applies to the FP32 paths without forward-backward averaging");
  if (ds.cov_mfma < 0 || ds.cov_mfma > 4) return fail(hipErrorInvalidValue, "bad cov_mfma");
  if (ds.cov_mfma == 1 && (ds.precision == COS_PREC_FP64 || ds.nsteer > 0 || ds.fb_avg || ds.z_mode ||
                           ds.p_mode || ds.refine))
    return fail(hipErrorInvalidValue, "the MFMA covariance takes an FP32 path and the default options");
  if ((ds.cov_mfma == 2 || ds.cov_mfma == 3) && ds.precision == COS_PREC_FP64)
    return fail(hipErrorInvalidValue, "the fused MFMA covariance takes an FP32 path");
  return 0;
}

// Whether the FP64 fallback pass can find work.  With one loading factor for the frame,
// cond(R_L) <= 1 + L/d; when that is below the threshold no pixel gets the condition
// flag, and when (1 + L/d) L 2^-24 < 0.01 an FP32 Cholesky cannot meet a non-positive
// pivot either (breakdown needs cond ~ 1/(L eps)), so the pass would process no pixel.
bool fallback_needed(const cos_capon_desc& ds) {
  if (ds.precision != COS_PREC_FP32_FALLBACK) return false;
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  if (tables || !(ds.d > 0.f)) return true;
  const double bound = 1.0 + (double)ds.L / (double)ds.d;
  const double thresh = ds.cond_thresh > 0 ? ds.cond_thresh : 1.0e6;
  return bound > thresh || bound * ds.L * 5.9604645e-8 >= 0.01;
}

int enqueue_frame(const FrameCtx& ctx, const cos_capon_desc& ds, const cos_cf32* x, float* P, cos_cf32* z,
                  cos_cf32* w, int wstride, cos_cf32* R, unsigned char* flags, hipStream_t s, bool fb_needed) {
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  const int Lmax = tables ? ds.Lmax : ds.L;
  const int Mmax = tables ? ds.Mmax : ds.M;
  cosk::CaponLaunch a = base_launch(ds, x);
  a.P = P;
  a.z = (float2*)z;
  a.w = (float2*)w;
  a.wstride = wstride;
  a.R = (float2*)R;
  a.Rstride = ds.L * ds.L;
  a.refine = ds.refine ? 1 : 0;
  a.mfma_gram = ds.cov_mfma == 2 ? 1 : ds.cov_mfma == 3 ? 2 : ds.cov_mfma == 4 ? 3 : 0;
  const char* why = nullptr;
  hipError_t e;
  const int Nzp = ds.Nz - 2 * ds.Yavg;
  if (tables && (e = cosk::plan_rows(ds.row_L, Nzp, ds.Yavg, ctx.bu.bk, ctx.bu.nb, ctx.rows + 8, ctx.rows, s)) !=
                    hipSuccess)
    return fail_hip(e, "row planning");
  // bucket b's launch: its device row list, a grid for every row (blocks past the count exit)
  auto bucket_launch = [&](cosk::CaponLaunch& c, int b) {
    c.rlist = ctx.rows + 8 + (size_t)b * Nzp;
    c.rcount = ctx.rows + b;
    c.npix = ds.Nx * Nzp;
    c.Lcap = ctx.bu.cap[b];
  };
  if (ds.cov_mfma == 1) {
    // covariance on the matrix cores into R (the caller's, or the plan's; Lmax^2 per pixel),
    // then the solve reads it back; one launch per row run, with one or four 32 x 32 tiles
    cosk::CaponLaunch c = a;
    c.R = R ? (float2*)R : ctx.R;
    const int Rs = Lmax * Lmax;
    if (tables) {
      for (int b = 0; b < ctx.bu.nb; ++b) {
        bucket_launch(c, b);
        if ((e = cosk::launch_cov_mfma(c, ctx.bu.cap[b] > 32 ? 2 : 1, Mmax, Rs, s)) != hipSuccess)
          return fail_hip(e, "MFMA covariance");
      }
    } else if ((e = cosk::launch_cov_mfma(c, Lmax > 32 ? 2 : 1, Mmax, Rs, s)) != hipSuccess) {
      return fail_hip(e, "MFMA covariance");
    }
    a.Rin = c.R;
    a.R = nullptr;
    a.Rstride = Rs;
  }
  if (ds.precision == COS_PREC_FP64) {
    if (tables) {
      for (int b = 0; b < ctx.bu.nb; ++b) {
        cosk::CaponLaunch c = a;
        bucket_launch(c, b);
        e = cosk::launch_capon_f64(c, ctx.bu.cap[b], Mmax, s, &why);
        if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "FP64 kernel");
      }
      return 0;
    }
    e = cosk::launch_capon_f64(a, Lmax, Mmax, s, &why);
    if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "FP64 kernel");
    return 0;
  }
  const bool fb = fb_needed;
  uint8_t* fl = (uint8_t*)flags;
  if (fb && !fl) fl = ctx.flags;
  a.flags = fl;
  if (tables) {
    // each lane shape over the rows of its bucket
    for (int b = 0; b < ctx.bu.nb; ++b) {
      cosk::CaponLaunch c = a;
      bucket_launch(c, b);
      e = cosk::launch_capon_f32_shape(c, ctx.bu.shape[b], Mmax, s, &why);
      if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "FP32 kernel");
    }
  } else {
    e = cosk::launch_capon_f32_shape(a, shape_for(ctx.shapes, Lmax), Mmax, s, &why);
    if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "FP32 kernel");
  }
  if (!fb) return 0;
  // With row tables the flagged pixels go to one list per FP64 lane shape (L <= 16, 32, 64)
  // by their row's L, so short rows run four to a wave instead of one.
  const int nfb = fb_lists(ds);
  cosk::RowBuckets fbk{};
  for (int L = 0; L <= 64; ++L) fbk.b[L] = (unsigned char)std::min(nfb - 1, L <= 16 ? 0 : L <= 32 ? 1 : 2);
  const int fcap[3] = {16, 32, 64};
  e = nfb > 1 ? cosk::compact_flags_rows(fl, a.npix, ctx.list, ctx.count, ds.row_L, ds.Nz - 2 * ds.Yavg, ds.Yavg,
                                         fbk, nfb, a.npix, s)
              : cosk::compact_flags(fl, a.npix, ctx.list, ctx.count, s);
  if (e != hipSuccess) return fail_hip(e, "compaction");
  // The FP64 pass over each device-side list: the entries the host expects (the bound's
  // count plus a slack for FP32 breakdowns) one block per group, the out-of-line
  // persistent kernel only over entries past them (none in the usual frame).  The split
  // changes only the launch shapes: every list entry is covered whatever the count.
  constexpr long kSlack = 4096;
  const int head = (int)std::min<long>(a.npix, ctx.fb_bound + kSlack);
  for (int b = 0; b < nfb; ++b) {
    const int cap = nfb > 1 ? fcap[b] : Lmax;
    cosk::CaponLaunch f = a;
    f.flags = nullptr;
    f.plist = ctx.list + (size_t)b * a.npix;
    f.pcount = ctx.count + b;
    f.pbase = 0;
    f.npix = head;
    f.Lcap = cap;
    e = cosk::launch_capon_f64_list(f, cap, Mmax, false, s, &why);
    if (e == hipSuccess && head < a.npix) {
      f.plist += head;
      f.pbase = head;
      f.npix = a.npix - head;
      e = cosk::launch_capon_f64_list(f, cap, Mmax, true, s, &why);
    }
    if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "FP64 fallback");
  }
  return 0;
}

bool same_desc(const cos_capon_desc& a, const cos_capon_desc& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

// FP64 fallback lists of a frame: one per lane shape up to the frame's largest L with row
// tables, one otherwise.
int fb_lists(const cos_capon_desc& ds) {
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  if (!tables) return 1;
  return ds.Lmax <= 16 ? 1 : ds.Lmax <= 32 ? 2 : 3;
}

}  // namespace

extern "C" {

int cos_capon_run(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x, float* P,
                  cos_cf32* z, cos_cf32* w, int wstride, cos_cf32* R, unsigned char* flags,
                  void* stream) {
  g_last_error.clear();
  if (!plan || !desc || !x) return fail(hipErrorInvalidValue, "plan, desc and x must be non-NULL");
  cos_capon_desc ds = *desc;
  if (int e = validate(ds, R, w, wstride)) return e;
  hipStream_t s = (hipStream_t)stream;
  const bool fb = fallback_needed(ds);
  const int npix = ds.Nx * (ds.Nz - 2 * ds.Yavg);
  hipError_t e;
  if (ds.cov_mfma == 1 && !R) {
    const bool tab = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
    const size_t Lm = tab ? ds.Lmax : ds.L;
    if ((e = plan->R.ensure((size_t)npix * Lm * Lm * sizeof(cos_cf32))) != hipSuccess)
      return fail_hip(e, "covariance buffer");
  }
  if (fb) {
    if (!flags && (e = plan->flags.ensure(npix)) != hipSuccess) return fail_hip(e, "flags buffer");
    const int nfb = fb_lists(ds);
    if ((e = plan->list.ensure(sizeof(int) * (size_t)npix * nfb)) != hipSuccess) return fail_hip(e, "list buffer");
    if ((e = plan->count.ensure(sizeof(int) * nfb)) != hipSuccess) return fail_hip(e, "count buffer");
    plan->fb_lists = nfb;
  }
  plan->last_stream = s;
  plan->fallback_ran = fb;
  const void* ptr[7] = {x, P, z, w, R, flags, nullptr};
  FrameGraph& g = plan->fg;
  const int gen = g_tune_gen.load();
  const bool replay = ds.graph && g.valid && same_desc(g.ds, ds) && std::memcmp(g.ptr, ptr, sizeof(ptr)) == 0 &&
                      g.wstride == wstride && g.tune_gen == gen && g.buf_gen == plan->buf_gen;
  FrameCtx ctx;
  ctx.shapes = tuned_shapes(ds);
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  if (tables) {
    ctx.bu = make_buckets(ds, ctx.shapes);
    const size_t Nzp = ds.Nz - 2 * ds.Yavg;
    if ((e = plan->rows.ensure(sizeof(int) * (8 + Nzp * ctx.bu.nb))) != hipSuccess) return fail_hip(e, "row lists");
    // the bound's count is not known without the tables: the FP64 head covers a sixteenth
    // of the pixels plus a slack, the persistent grid the rest
    ctx.fb_bound = npix / 16;
  } else {
    const float th = (float)(ds.cond_thresh > 0 ? ds.cond_thresh : 1.0e6);
    ctx.fb_bound = (1.f + (float)ds.L / ds.d) > th ? npix : 0;
  }
  ctx.rows = plan->rows.as<int>();
  ctx.flags = plan->flags.as<uint8_t>();
  ctx.list = plan->list.as<int>();
  ctx.count = plan->count.as<int>();
  ctx.R = plan->R.as<float2>();
  if (!ds.graph) return enqueue_frame(ctx, ds, x, P, z, w, wstride, R, flags, s, fb);

  if (!replay) {
    g.reset();
    if (!plan->capture_stream && (e = hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking)) != hipSuccess)
      return fail_hip(e, "capture stream");
    if ((e = hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return fail_hip(e, "begin capture");
    const int r = enqueue_frame(ctx, ds, x, P, z, w, wstride, R, flags, plan->capture_stream, fb);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(plan->capture_stream, &graph);
    if (r) {
      if (graph) (void)hipGraphDestroy(graph);
      return r;
    }
    if (e != hipSuccess) return fail_hip(e, "end capture");
    g.graph = graph;
    if ((e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
      g.reset();
      return fail_hip(e, "graph instantiate");
    }
    g.ds = ds;
    std::memcpy(g.ptr, ptr, sizeof(ptr));
    g.wstride = wstride;
    g.tune_gen = gen;
    g.buf_gen = plan->buf_gen;
    g.valid = true;
  }
  if ((e = hipGraphLaunch(g.exec, s)) != hipSuccess) return fail_hip(e, "graph launch");
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Entry point without allocation: the FP64 fallback's scratch (per-pixel flags when the
// caller passes none, the compacted pixel list, its device-side length) lives in one
// file-scope device array of fixed size instead of plan buffers.  Each device holds its
// own copy until the process ends; host code uses only the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in).  The count
// is cleared by a fill kernel at the start of every compaction, so no call depends on
// what an earlier call left behind; calls that overlap on one device share the array and
// race.
// ---------------------------------------------------------------------------------------
namespace {

constexpr size_t kScratchCountBytes = 256;   // the list length, padded to the array's alignment
constexpr size_t kScratchBytes =
    kScratchCountBytes + (size_t)COS_CAPON_NOALLOC_MAX_PIXELS * (sizeof(int) + sizeof(uint8_t));

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's scratch array, looked up once per device.
hipError_t scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" {

int cos_capon_run_noalloc(const cos_capon_desc* desc, const cos_cf32* x, float* P, cos_cf32* z, cos_cf32* w,
                          int wstride, unsigned char* flags, void* stream) {
  g_last_error.clear();
  if (!desc || !x) return fail(hipErrorInvalidValue, "desc and x must be non-NULL");
  const cos_capon_desc ds = *desc;
  if (ds.row_L || ds.row_M || ds.row_off || ds.row_d)
    return fail(hipErrorInvalidValue, "the no-allocation entry point takes a uniform L, M and d (no row tables)");
  if (ds.graph) return fail(hipErrorInvalidValue, "the no-allocation entry point does not capture graphs");
  if (ds.cov_mfma) return fail(hipErrorInvalidValue, "the no-allocation entry point has no matrix-core covariance");
  if (int e = validate(ds, nullptr, w, wstride)) return e;
  const long npix = (long)ds.Nx * (ds.Nz - 2 * ds.Yavg);
  if (npix > COS_CAPON_NOALLOC_MAX_PIXELS)
    return fail(hipErrorInvalidValue, "the no-allocation entry point takes at most COS_CAPON_NOALLOC_MAX_PIXELS "
                                      "(2097152) output pixels per call");
  const bool fb = fallback_needed(ds);
  FrameCtx ctx;
  ctx.shapes = tuned_shapes(ds);
  const float th = (float)(ds.cond_thresh > 0 ? ds.cond_thresh : 1.0e6);
  ctx.fb_bound = (1.f + (float)ds.L / ds.d) > th ? npix : 0;
  if (fb) {
    void* base = nullptr;
    if (hipError_t e = scratch_address(&base)) return fail_hip(e, "scratch address");
    unsigned char* b = static_cast<unsigned char*>(base);
    ctx.count = reinterpret_cast<int*>(b);
    ctx.list = reinterpret_cast<int*>(b + kScratchCountBytes);
    ctx.flags = b + kScratchCountBytes + (size_t)COS_CAPON_NOALLOC_MAX_PIXELS * sizeof(int);
  }
  return enqueue_frame(ctx, ds, x, P, z, w, wstride, nullptr, flags, (hipStream_t)stream, fb);
}

int cos_capon_autotune(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x, float* P,
                       cos_cf32* z, cos_cf32* w, int wstride, void* stream, int* shapes_out, float* ms_out) {
  g_last_error.clear();
  if (!plan || !desc || !x) return fail(hipErrorInvalidValue, "plan, desc and x must be non-NULL");
  cos_capon_desc ds = *desc;
  if (ds.precision == COS_PREC_FP64) return fail(hipErrorInvalidValue, "the autotuner times the FP32 kernels");
  if (ds.nsteer > 0) return fail(hipErrorInvalidValue, "the autotuner takes unsteered frames");
  if (int e = validate(ds, nullptr, w, wstride)) return e;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  const int Nzp = ds.Nz - 2 * ds.Yavg;
  constexpr int kRows = 128;   // output rows (all beams) timed per subarray length
  // one run of rows per distinct L: the first rows with that length, up to kRows
  std::map<int, std::pair<int, int>> runs;
  if (tables) {
    std::vector<int> hL(ds.Nz);
    if ((e = hipMemcpyAsync(hL.data(), ds.row_L, sizeof(int) * ds.Nz, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return fail_hip(e, "read L table");
    for (int jj = 0; jj < Nzp;) {
      const int Lr = hL[jj + ds.Yavg];
      int j2 = jj + 1;
      while (j2 < Nzp && j2 - jj < kRows && hL[j2 + ds.Yavg] == Lr) ++j2;
      if (Lr >= 1 && Lr <= ds.Lmax && !runs.count(Lr)) runs[Lr] = {jj, j2 - jj};
      jj = j2;
    }
  } else {
    runs[ds.L] = {0, Nzp < kRows ? Nzp : kRows};
  }
  const int Mmax = tables ? ds.Mmax : ds.M;
  if (shapes_out) std::fill(shapes_out, shapes_out + 65, 0);
  if (ms_out) std::fill(ms_out, ms_out + 65 * cosk::kShapes, -1.f);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail_hip(e, "tuning events");
  }
  TuneTable t{};
  int rc = 0;
  const char* why = nullptr;
  for (const auto& kv : runs) {
    const int Lr = kv.first;
    float best_ms = 0.f;
    for (int sh = 1; sh <= cosk::kShapes && !rc; ++sh) {
      if (!cosk::shape_fits(sh, Lr)) continue;
      cosk::CaponLaunch a = base_launch(ds, x);
      a.P = P;
      a.z = (float2*)z;
      a.w = (float2*)w;
      a.wstride = wstride;
      a.refine = ds.refine ? 1 : 0;
      a.jlo = kv.second.first;
      a.jn = kv.second.second;
      a.npix = ds.Nx * a.jn;
      constexpr int kReps = 3;
      if ((e = cosk::launch_capon_f32_shape(a, sh, Mmax, s, &why)) != hipSuccess ||   // warm-up
          (e = hipEventRecord(e0, s)) != hipSuccess) {
        rc = why ? fail(e, why) : fail_hip(e, "tuning launch");
        break;
      }
      for (int r = 0; r < kReps && e == hipSuccess; ++r) e = cosk::launch_capon_f32_shape(a, sh, Mmax, s, &why);
      float ms = 0.f;
      if (e != hipSuccess || (e = hipEventRecord(e1, s)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess ||
          (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) {
        rc = why ? fail(e, why) : fail_hip(e, "tuning launch");
        break;
      }
      ms /= kReps;
      if (ms_out) ms_out[Lr * cosk::kShapes + sh - 1] = ms;
      if (!t[Lr] || ms < best_ms) {
        t[Lr] = (unsigned char)sh;
        best_ms = ms;
      }
    }
    if (rc) break;
    if (shapes_out) shapes_out[Lr] = t[Lr];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneTable& dst = g_tune[tune_key(ds)];
    for (int Lr = 1; Lr <= 64; ++Lr)
      if (t[Lr]) dst[Lr] = t[Lr];
  }
  ++g_tune_gen;
  return 0;
}

int cos_capon_tuning_set(const cos_capon_desc* desc, const int* shapes) {
  g_last_error.clear();
  if (!desc || !shapes) return fail(hipErrorInvalidValue, "desc and shapes must be non-NULL");
  TuneTable t{};
  for (int Lr = 1; Lr <= 64; ++Lr) {
    const int sh = shapes[Lr];
    if (sh < 0 || sh > cosk::kShapes || (sh && !cosk::shape_fits(sh, Lr)))
      return fail(hipErrorInvalidValue, "shape table entry does not hold its subarray length");
    t[Lr] = (unsigned char)sh;
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[tune_key(*desc)] = t;
  }
  ++g_tune_gen;
  return 0;
}

int cos_capon_subspace_run(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* V, int Nb,
                           const cos_cf32* x, float* P, cos_cf32* z, cos_cf32* w, void* stream) {
  g_last_error.clear();
  if (!plan || !desc || !x || !V) return fail(hipErrorInvalidValue, "plan, desc, V and x must be non-NULL");
  cos_capon_desc ds = *desc;
  if (ds.row_L || ds.row_M || ds.row_off || ds.row_d || ds.nsteer)
    return fail(hipErrorInvalidValue, "subspace path takes uniform L, M, d and no steering");
  if (ds.z_mode || ds.p_mode) return fail(hipErrorInvalidValue, "subspace path takes the default z and P");
  if (int e = check_args(ds.d, ds.L, ds.Yavg, ds.M, ds.Nx, ds.Nz, Nb)) return e;
  if (Nb < 1) return fail(hipErrorInvalidValue, "need 1 <= Nb <= L");
  if (ds.off < 0 || ds.off + ds.M > ds.Mstride) return fail(hipErrorInvalidValue, "off + M > Mstride");
  const int L = ds.L;
  const size_t lds = ((size_t)L * Nb + (size_t)Nb * Nb + 3 * (size_t)Nb + L) * 16 + (size_t)L * L * 8;
  if (lds > 64 * 1024) return fail(hipErrorInvalidValue, "subspace path: L and Nb too large for one block");
  const size_t npx = (size_t)ds.Nx * (ds.Nz - 2 * ds.Yavg);
  hipError_t e;
  if ((e = plan->R.ensure(npx * L * L * sizeof(cos_cf32))) != hipSuccess) return fail_hip(e, "covariance buffer");
  hipStream_t s = (hipStream_t)stream;
  // unloaded, unnormalised covariance of every pixel (FP32), then the subspace solve (FP64)
  ds.d = 0.f;
  cosk::CaponLaunch a = base_launch(ds, x);
  a.R = plan->R.as<float2>();
  a.Rstride = L * L;
  a.cov_only = 1;
  const char* why = nullptr;
  if ((e = cosk::launch_capon_f32(a, L, ds.M, s, &why)) != hipSuccess)
    return why ? fail(e, why) : fail_hip(e, "covariance kernel");
  if ((e = cosk::subspace_capon(plan->R.as<float2>(), (const float2*)V, (const float2*)x, L, Nb, ds.M, ds.Mstride,
                                ds.off, ds.Yavg, ds.Nx, ds.Nz, desc->d, (float2*)w, (float2*)z, P, s)) != hipSuccess)
    return fail_hip(e, "subspace kernel");
  return 0;
}

namespace {

int tdib_common(const cos_cf32* mfdata, int n_h, int n_t, double t0, double dt, const double* taxis,
                const double* x_r, const double* D, double c, double fc, const double* xarr, const double* yarr,
                int n_x, int n_y, float fill, cos_cf32* cube, cos_cf32* image, void* stream) {
  g_last_error.clear();
  if (!mfdata || !x_r || !D || !xarr || !yarr) return fail(hipErrorInvalidValue, "NULL input");
  if (n_h < 1 || n_t < 1 || n_x < 1 || n_y < 1) return fail(hipErrorInvalidValue, "sizes must be >= 1");
  if (!taxis && (!(dt > 0.0) || !std::isfinite(t0)))
    return fail(hipErrorInvalidValue, "need dt > 0 and a finite t0");
  if (!(c > 0.0) || !std::isfinite(fc)) return fail(hipErrorInvalidValue, "need c > 0 and a finite fc");
  if ((long)n_x * n_y > (1L << 31) - 1) return fail(hipErrorInvalidValue, "too many pixels for one call");
  cosk::TdibLaunch a;
  std::memset(&a, 0, sizeof(a));
  a.mf = (const float2*)mfdata;
  a.n_h = n_h;
  a.n_t = n_t;
  a.t0 = t0;
  a.dt = taxis ? 1.0 : dt;
  a.taxis = taxis;
  a.x_r = x_r;
  a.D = D;
  a.inv_c = 1.0 / c;
  a.fc = fc;
  a.xarr = xarr;
  a.yarr = yarr;
  a.n_x = n_x;
  a.n_y = n_y;
  a.fill = fill;
  a.cube = (float2*)cube;
  a.image = (float2*)image;
  hipError_t e = cosk::launch_tdib(a, (hipStream_t)stream);
  return e == hipSuccess ? 0 : fail_hip(e, "tdib kernel");
}

}  // namespace

int cos_tdib_run(const cos_cf32* mfdata, int n_h, int n_t, double t0, double dt, const double* x_r,
                 const double* D, double c, double fc, const double* xarr, const double* yarr,
                 int n_x, int n_y, float fill, cos_cf32* cube, cos_cf32* image, void* stream) {
  return tdib_common(mfdata, n_h, n_t, t0, dt, nullptr, x_r, D, c, fc, xarr, yarr, n_x, n_y, fill, cube, image,
                     stream);
}

int cos_tdib_run_axis(const cos_cf32* mfdata, int n_h, int n_t, const double* taxis, const double* x_r,
                      const double* D, double c, double fc, const double* xarr, const double* yarr,
                      int n_x, int n_y, float fill, cos_cf32* cube, cos_cf32* image, void* stream) {
  if (!taxis) {
    g_last_error = "taxis is NULL";
    return hipErrorInvalidValue;
  }
  return tdib_common(mfdata, n_h, n_t, 0.0, 0.0, taxis, x_r, D, c, fc, xarr, yarr, n_x, n_y, fill, cube, image,
                     stream);
}

int cos_lca_run(const cos_cf32* Z, int Nx, int Ny, int Nw, const float* wrf, int wy, int wx, int mode,
                cos_cf32* img, int* win, void* stream) {
  g_last_error.clear();
  if (!Z || !wrf) return fail(hipErrorInvalidValue, "NULL input");
  if (Nx < 1 || Ny < 1 || Nw < 1 || wy < 1 || wx < 1) return fail(hipErrorInvalidValue, "sizes must be >= 1");
  if (wy % 2 == 0 || wx % 2 == 0) return fail(hipErrorInvalidValue, "the wrf filter must have odd sizes");
  if (mode != COS_LCA_REFERENCE && mode != COS_LCA_CENTRED) return fail(hipErrorInvalidValue, "bad LCA mode");
  if ((long)Nx * Ny > (1L << 31) - 1 || (long)Nx * Ny * Nw > (1L << 40))
    return fail(hipErrorInvalidValue, "too many pixels for one call");
  cosk::LcaLaunch a;
  std::memset(&a, 0, sizeof(a));
  a.Z = (const float2*)Z;
  a.Nx = Nx;
  a.Ny = Ny;
  a.Nw = Nw;
  a.wrf = wrf;
  a.wy = wy;
  a.wx = wx;
  a.img = (float2*)img;
  a.win = win;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (mode == COS_LCA_REFERENCE) {
    // the reference picks its branch from the filter's values (lcacx.pyx:94-100,122,139,185)
    if ((long)wy * wx > 4096) return fail(hipErrorInvalidValue, "reference mode takes filters of at most 4096 taps");
    float h[4096];
    if ((e = hipMemcpyAsync(h, wrf, sizeof(float) * wy * wx, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return fail_hip(e, "read wrf");
    bool ones = wx == 1;
    for (int i = 0; i < wy * wx && ones; ++i) ones = h[i] == 1.0f;
    const int Ym = wy / 2, Xm = wx / 2;
    if (wy == 1 && wx == 1 && h[0] != 0.0f) {
      a.branch = 1;
    } else if (ones) {
      if (Ny <= wy)
        return fail(hipErrorInvalidValue, "reference mode: the all-ones range filter needs Ny > its length "
                                          "(the reference writes row Y_AVG outside the image)");
      if (Nw > 256) return fail(hipErrorInvalidValue, "reference mode: the all-ones range filter takes Nw <= 256");
      a.branch = 2;
      a.rows_per_chunk = std::max(1, std::min(1024, 8192 / Nw));   // up to 64 KB of FP64 sums in LDS
    } else {
      if (Ny - 2 * Ym > 0 && Nx - 2 * Xm > 0 && Ny > Nx)
        return fail(hipErrorInvalidValue, "reference mode: this filter needs Ny <= Nx (the reference's general "
                                          "branch reads past its |Z|^2 array on taller images)");
      a.branch = 3;
    }
  }
  e = cosk::launch_lca(a, s);
  return e == hipSuccess ? 0 : fail_hip(e, "lca kernel");
}

int cos_capon_beamspace_run(cos_capon_plan* plan, const cos_capon_desc* desc, int Nb, const cos_cf32* x,
                            float* P, cos_cf32* z, cos_cf32* w_bs, cos_cf32* w, int wstride, void* stream) {
  g_last_error.clear();
  if (!plan || !desc || !x) return fail(hipErrorInvalidValue, "plan, desc and x must be non-NULL");
  const cos_capon_desc& ds = *desc;
  if (ds.row_L || ds.row_M || ds.row_off || ds.row_d || ds.nsteer)
    return fail(hipErrorInvalidValue, "beamspace path takes uniform L, M, d and no steering");
  if (ds.fb_avg || ds.z_mode || ds.p_mode)
    return fail(hipErrorInvalidValue, "beamspace path takes the default options");
  if (int e = check_args(ds.d, ds.L, ds.Yavg, ds.M, ds.Nx, ds.Nz, Nb)) return e;
  if (Nb < 1 || Nb > 8) return fail(hipErrorInvalidValue, "beamspace path supports 1 <= Nb <= 8");
  if (ds.L > 64) return fail(hipErrorInvalidValue, "beamspace path supports L <= 64");
  if (ds.off < 0 || ds.off + ds.M > ds.Mstride) return fail(hipErrorInvalidValue, "off + M > Mstride");
  if (2 * ds.Yavg > 120) return fail(hipErrorInvalidValue, "beamspace path supports Yavg <= 60");
  // the kernel addresses a wave's 64 rows with 32-bit buffer offsets below 2^31
  if ((long long)ds.Mstride * 64 * 8 >= (1LL << 31))
    return fail(hipErrorInvalidValue, "beamspace path supports Mstride < 4194304");
  if (w && wstride < ds.L) return fail(hipErrorInvalidValue, "wstride < L");
  hipError_t e;
  if (plan->bs_Nb != Nb || plan->bs_L != ds.L) {
    // beams beta(b) (RealTimeCapon/capon_kernel.cu:53-60 without the never-firing even-Nb +1)
    std::vector<float> h(2 * ((size_t)Nb * ds.L + Nb));
    const int L = ds.L;
    for (int b = 0; b < Nb; ++b) {
      const int beta = b > Nb / 2 ? b + L - 2 * (Nb / 2) - 1 : b;
      for (int l = 0; l < L; ++l) {
        const double ph = -2.0 * M_PI * (double)((long)beta * l % L) / (double)L;
        h[2 * ((size_t)b * L + l)] = (float)std::cos(ph);
        h[2 * ((size_t)b * L + l) + 1] = (float)std::sin(ph);
      }
      const double ph = 2.0 * M_PI * (double)beta / (double)L;
      h[2 * ((size_t)Nb * L + b)] = (float)std::cos(ph);
      h[2 * ((size_t)Nb * L + b) + 1] = (float)std::sin(ph);
    }
    if ((e = plan->bs_tw.ensure(h.size() * sizeof(float))) != hipSuccess) return fail_hip(e, "twiddle buffer");
    if ((e = hipMemcpy(plan->bs_tw.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
      return fail_hip(e, "copy twiddles");
    plan->bs_Nb = Nb;
    plan->bs_L = ds.L;
  }
  cosk::BeamspaceLaunch a;
  std::memset(&a, 0, sizeof(a));
  a.x = (const float2*)x;
  a.Nx = ds.Nx; a.Nz = ds.Nz; a.Mstride = ds.Mstride; a.M = ds.M; a.L = ds.L; a.off = ds.off;
  a.Y = ds.Yavg; a.Nb = Nb; a.d = ds.d;
  a.rsqrtL = (float)(1.0 / std::sqrt((double)ds.L));
  a.sqrtL = (float)std::sqrt((double)ds.L);
  a.tw = plan->bs_tw.as<float2>();
  a.omega = plan->bs_tw.as<float2>() + (size_t)Nb * ds.L;
  a.P = P; a.z = (float2*)z; a.wbs = (float2*)w_bs; a.w = (float2*)w; a.wstride = wstride;
  if ((e = cosk::launch_beamspace(a, (hipStream_t)stream)) != hipSuccess) return fail_hip(e, "beamspace kernel");
  return 0;
}

int cos_capon_backward(cos_capon_plan* plan, const cos_capon_desc* desc, const cos_cf32* x,
                       const float* grad_P, cos_cf32* grad_x, float* grad_d, void* stream) {
  g_last_error.clear();
  if (!plan || !desc || !x || !grad_P || !grad_x) return fail(hipErrorInvalidValue, "NULL argument");
  cos_capon_desc ds = *desc;
  ds.graph = 0;
  if (ds.nsteer) return fail(hipErrorInvalidValue, "the adjoint does not take steering vectors");
  if (ds.fb_avg || ds.z_mode || ds.p_mode)
    return fail(hipErrorInvalidValue, "the adjoint is defined for the default P (no averaging options)");
  if (int e = validate(ds, nullptr, nullptr, 0)) return e;
  const bool tables = ds.row_L || ds.row_M || ds.row_off || ds.row_d;
  const int Lmax = tables ? ds.Lmax : ds.L;
  const int Mmax = tables ? ds.Mmax : ds.M;
  cosk::CaponLaunch a = base_launch(ds, x);
  a.gP = grad_P;
  a.gx = (float2*)grad_x;
  a.gd = grad_d;
  const char* why = nullptr;
  hipError_t e = cosk::launch_capon_bwd(a, Lmax, Mmax, (hipStream_t)stream, &why);
  if (e != hipSuccess) return why ? fail(e, why) : fail_hip(e, "adjoint kernel");
  return 0;
}

int cos_capon_last_fallback_count(cos_capon_plan* plan, int* count) {
  if (!plan || !count) return fail(hipErrorInvalidValue, "NULL argument");
  *count = 0;
  if (!plan->fallback_ran) return 0;
  hipError_t e = hipStreamSynchronize(plan->last_stream);
  if (e != hipSuccess) return fail_hip(e, "synchronize");
  int c[3] = {0, 0, 0};
  e = hipMemcpy(c, plan->count.p, sizeof(int) * plan->fb_lists, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail_hip(e, "copy count");
  *count = c[0] + c[1] + c[2];
  return 0;
}

int cos_capon_get_capon_p(cos_cf32* z, cos_cf32* w, cos_cf32* R, float* P, const cos_cf32* x, float d,
                          int L, int Yavg, int M, int Nx, int Nz, int Nb) {
  g_last_error.clear();
  if (int e = check_args(d, L, Yavg, M, Nx, Nz, Nb)) return e;
  if (!x) return fail(hipErrorInvalidValue, "x is NULL");
  std::lock_guard<std::mutex> lk(g_default_mu);
  cos_capon_plan* pl = default_plan();
  const size_t npx = (size_t)Nx * (Nz - 2 * Yavg);
  const size_t xb = (size_t)Nx * Nz * M * sizeof(cos_cf32);
  hipError_t e;
  const bool needR = R || Nb > 0;
  if ((e = pl->x.ensure(xb)) != hipSuccess) return fail_hip(e, "device buffer x");
  if (z && (e = pl->z.ensure(npx * sizeof(cos_cf32))) != hipSuccess) return fail_hip(e, "device buffer z");
  if (w && (e = pl->w.ensure(npx * L * sizeof(cos_cf32))) != hipSuccess) return fail_hip(e, "device buffer w");
  if (P && (e = pl->P.ensure(npx * sizeof(float))) != hipSuccess) return fail_hip(e, "device buffer P");
  if (needR && (e = pl->R.ensure(npx * L * L * sizeof(cos_cf32))) != hipSuccess)
    return fail_hip(e, "device buffer R");
  if ((e = hipMemcpy(pl->x.p, x, xb, hipMemcpyHostToDevice)) != hipSuccess) return fail_hip(e, "copy x");

  cos_capon_desc ds;
  std::memset(&ds, 0, sizeof(ds));
  ds.Nx = Nx; ds.Nz = Nz; ds.Mstride = M; ds.M = M; ds.L = L; ds.off = 0; ds.Yavg = Yavg; ds.d = d;
  ds.precision = COS_PREC_FP32;
  if (Nb == 0 && L > 64) {
    // subarrays past the fused kernel's lane shapes: one block per pixel in LDS
    cosk::CaponLaunch a = base_launch(ds, pl->x.p);
    a.P = P ? pl->P.as<float>() : nullptr;
    a.z = z ? pl->z.as<float2>() : nullptr;
    a.w = w ? pl->w.as<float2>() : nullptr;
    a.wstride = L;
    a.R = R ? pl->R.as<float2>() : nullptr;
    a.Rstride = L * L;
    if ((e = cosk::launch_capon_wide(a, nullptr)) != hipSuccess) return fail_hip(e, "wide kernel");
  } else if (Nb == 0) {
    int r = cos_capon_run(pl, &ds, pl->x.as<cos_cf32>(), P ? pl->P.as<float>() : nullptr,
                          z ? pl->z.as<cos_cf32>() : nullptr, w ? pl->w.as<cos_cf32>() : nullptr, L,
                          R ? pl->R.as<cos_cf32>() : nullptr, nullptr, nullptr);
    if (r) return r;
  } else {
    // Reference default beamspace: element-space loaded S_L, then B S_L B^H (Capon.cpp:285-402).
    cosk::CaponLaunch a = base_launch(ds, pl->x.p);
    a.R = pl->R.as<float2>();
    a.Rstride = L * L;
    a.cov_only = 1;
    const char* why = nullptr;
    if ((e = cosk::launch_capon_f32(a, L, M, nullptr, &why)) != hipSuccess)
      return why ? fail(e, why) : fail_hip(e, "covariance kernel");
    std::vector<float> Bh = butler_host(Nb, L);
    if ((e = pl->B.ensure(Bh.size() * sizeof(float))) != hipSuccess) return fail_hip(e, "device buffer B");
    if ((e = hipMemcpy(pl->B.p, Bh.data(), Bh.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
      return fail_hip(e, "copy B");
    e = cosk::beamspace_matrix(pl->R.as<float2>(), pl->B.as<float2>(), pl->x.as<float2>(), L, Nb, M, Yavg, Nx,
                               Nz, w ? pl->w.as<float2>() : nullptr, z ? pl->z.as<float2>() : nullptr,
                               P ? pl->P.as<float>() : nullptr, nullptr);
    if (e != hipSuccess) return fail_hip(e, "beamspace kernel");
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_hip(e, "kernel execution");
  if (z && (e = hipMemcpy(z, pl->z.p, npx * sizeof(cos_cf32), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail_hip(e, "copy z");
  if (w && (e = hipMemcpy(w, pl->w.p, npx * L * sizeof(cos_cf32), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail_hip(e, "copy w");
  if (P && (e = hipMemcpy(P, pl->P.p, npx * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail_hip(e, "copy P");
  if (R && (e = hipMemcpy(R, pl->R.p, npx * L * L * sizeof(cos_cf32), hipMemcpyDeviceToHost)) != hipSuccess)
    return fail_hip(e, "copy R");
  return 0;
}

int cos_capon_get_capon(cos_cf32* z, cos_cf32* w, cos_cf32* R, const cos_cf32* x, float d, int L, int Yavg,
                        int M, int Nx, int Nz, int Nb) {
  return cos_capon_get_capon_p(z, w, R, nullptr, x, d, L, Yavg, M, Nx, Nz, Nb);
}

int cos_buildr_get_r(cos_cf32* R, const cos_cf32* x, float d, int L, int Yavg, int M, int Nx, int Nz, int Nb,
                     int R_on_gpu, int x_on_gpu) {
  g_last_error.clear();
  // Argument checks of RealTimeCapon/BuildR.cpp:37-48.
  if (Nx < 1 || Nz < 1 || L < 1 || M < 1 || Yavg < 0 || Nb < 0)
    return fail(hipErrorInvalidValue, "Nx, Nz, L, M, Yavg or Nb were <= 0");
  if (L > M) return fail(hipErrorInvalidValue, "L > M");
  if (2 * Yavg >= Nz) return fail(hipErrorInvalidValue, "Nz <= Yavg");
  if (!R || !x) return fail(hipErrorInvalidValue, "R and x must be non-NULL");
  if (Nb == 0 && L > 128) return fail(hipErrorInvalidValue, "L > 128 is not supported (the reference buildR's limit)");
  std::lock_guard<std::mutex> lk(g_default_mu);
  cos_capon_plan* pl = default_plan();
  const int K = M - L + 1;
  const size_t npx = (size_t)Nx * (Nz - 2 * Yavg);
  const int Ld = Nb > 0 ? Nb : L;
  const size_t Rb = npx * Ld * Ld * sizeof(cos_cf32);
  const size_t xb = (size_t)Nx * Nz * (Nb > 0 ? (size_t)K * Nb : (size_t)M) * sizeof(cos_cf32);
  hipError_t e;
  const cos_cf32* xd = x;
  cos_cf32* Rd = R;
  if (!x_on_gpu) {
    if ((e = pl->x.ensure(xb)) != hipSuccess) return fail_hip(e, "device buffer x");
    if ((e = hipMemcpy(pl->x.p, x, xb, hipMemcpyHostToDevice)) != hipSuccess) return fail_hip(e, "copy x");
    xd = pl->x.as<cos_cf32>();
  }
  if (!R_on_gpu) {
    if ((e = pl->R.ensure(Rb)) != hipSuccess) return fail_hip(e, "device buffer R");
    Rd = pl->R.as<cos_cf32>();
  }
  if (Nb == 0) {
    cos_capon_desc ds;
    std::memset(&ds, 0, sizeof(ds));
    ds.Nx = Nx; ds.Nz = Nz; ds.Mstride = M; ds.M = M; ds.L = L; ds.Yavg = Yavg; ds.d = d;
    cosk::CaponLaunch a = base_launch(ds, xd);
    a.R = (float2*)Rd;
    a.Rstride = L * L;
    a.cov_only = 1;
    const char* why = nullptr;
    if ((e = cosk::launch_capon_f32(a, L, M, nullptr, &why)) != hipSuccess)
      return why ? fail(e, why) : fail_hip(e, "covariance kernel");
  } else {
    if ((e = cosk::buildr_beamspace((const float2*)xd, (float2*)Rd, d, Nb, K, Yavg, Nx, Nz, nullptr)) != hipSuccess)
      return fail_hip(e, "beamspace covariance kernel");
  }
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_hip(e, "kernel execution");
  if (!R_on_gpu && (e = hipMemcpy(R, Rd, Rb, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "copy R");
  return 0;
}

int cos_solver_solve(cos_cf32* x, const cos_cf32* A, const cos_cf32* b, int N, int batch, int x_on_gpu,
                     int A_on_gpu, int b_on_gpu) {
  g_last_error.clear();
  // Solver::solve accepts 1 <= N <= 72 (RealTimeCapon/Solver.cpp:34-42).
  if (N < 1 || N > 72 || batch < 1) return fail(hipErrorInvalidValue, "need 1 <= N <= 72 and batch >= 1");
  if (!x || !A || !b) return fail(hipErrorInvalidValue, "x, A and b must be non-NULL");
  std::lock_guard<std::mutex> lk(g_default_mu);
  cos_capon_plan* pl = default_plan();
  const size_t Ab = (size_t)batch * N * N * sizeof(cos_cf32), vb = (size_t)batch * N * sizeof(cos_cf32);
  hipError_t e;
  const cos_cf32* Ad = A;
  const cos_cf32* bd = b;
  cos_cf32* xd = x;
  if (!A_on_gpu) {
    if ((e = pl->R.ensure(Ab)) != hipSuccess) return fail_hip(e, "device buffer A");
    if ((e = hipMemcpy(pl->R.p, A, Ab, hipMemcpyHostToDevice)) != hipSuccess) return fail_hip(e, "copy A");
    Ad = pl->R.as<cos_cf32>();
  }
  if (!b_on_gpu) {
    if ((e = pl->z.ensure(vb)) != hipSuccess) return fail_hip(e, "device buffer b");
    if ((e = hipMemcpy(pl->z.p, b, vb, hipMemcpyHostToDevice)) != hipSuccess) return fail_hip(e, "copy b");
    bd = pl->z.as<cos_cf32>();
  }
  if (!x_on_gpu) {
    if ((e = pl->w.ensure(vb)) != hipSuccess) return fail_hip(e, "device buffer x");
    xd = pl->w.as<cos_cf32>();
  }
  if ((e = cosk::solve_lu((const float2*)Ad, (const float2*)bd, (float2*)xd, N, batch, nullptr)) != hipSuccess)
    return fail_hip(e, "solver kernel");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return fail_hip(e, "kernel execution");
  if (!x_on_gpu && (e = hipMemcpy(x, xd, vb, hipMemcpyDeviceToHost)) != hipSuccess) return fail_hip(e, "copy x");
  return 0;
}

}  // extern "C"
