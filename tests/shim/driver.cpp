// Calls the library through the header-only shims include/Capon.h, BuildR.h and Solver.h
// the way the Cython code generated from RealTimeCapon/Capon.pxd, BuildR.pxd and
// Solver.pxd does: one object, every argument an lvalue bound to a reference parameter.
//
//   driver errors            argument errors only (no device call): prints "<case> <code> <message>"
//   driver run <dir>         a small frame through all three classes; writes x.bin (input
//                            cube [2][9][12]), z.bin, w.bin, R.bin, Rg.bin (getR), A.bin,
//                            b.bin, xs.bin (solve) as raw complex64
#include <Capon.h>
#include <BuildR.h>
#include <Solver.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

typedef Complex<float> cf;

static void put(const std::string& path, const std::vector<cf>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(v.data(), sizeof(cf), v.size(), f);
  std::fclose(f);
}

static void report(const char* name, int code) { std::printf("%s %d %s\n", name, code, cos_capon_last_error()); }

static int errors() {
  Capon capon;
  BuildR buildr;
  Solver solver;
  std::vector<cf> buf(64);
  cf *z = buf.data(), *w = buf.data(), *R = buf.data(), *x = buf.data();
  float d = 0.1f;
  int L = 9, Yavg = 1, M = 8, Nx = 2, Nz = 6, Nb = 0;
  report("capon_L_gt_M", capon.getCapon(z, w, R, x, d, L, Yavg, M, Nx, Nz, Nb));
  L = 4;
  Yavg = 3;
  report("capon_2Y_ge_Nz", capon.getCapon(z, w, R, x, d, L, Yavg, M, Nx, Nz, Nb));
  Yavg = 1;
  Nb = 5;
  report("capon_Nb_gt_L", capon.getCapon(z, w, R, x, d, L, Yavg, M, Nx, Nz, Nb));
  Nb = 0;
  bool Rg = false, xg = false;
  Yavg = 3;
  report("buildr_Nz", buildr.getR(R, x, d, L, Yavg, M, Nx, Nz, Nb, Rg, xg));
  int N = 73, batch = 1;
  bool g0 = false, g1 = false, g2 = false;
  report("solver_N", solver.solve(x, R, z, N, batch, g0, g1, g2));
  N = 4;
  batch = 0;
  report("solver_batch", solver.solve(x, R, z, N, batch, g0, g1, g2));
  return 0;
}

static int run(const std::string& dir) {
  int Nx = 2, Nz = 9, M = 12, L = 5, Yavg = 1, Nb = 0;
  float d = 0.2f;
  const int Nzp = Nz - 2 * Yavg;
  std::vector<cf> xv((size_t)Nx * Nz * M), zv((size_t)Nx * Nzp), wv((size_t)Nx * Nzp * L),
      Rv((size_t)Nx * Nzp * L * L), Rg((size_t)Nx * Nzp * L * L);
  for (size_t i = 0; i < xv.size(); ++i)
    xv[i] = cf((float)std::sin(0.37 * i + 0.1), (float)std::cos(0.91 * (double)(i * i % 17) + 0.3));
  cf *z = zv.data(), *w = wv.data(), *R = Rv.data(), *x = xv.data(), *R2 = Rg.data();
  Capon capon;
  int e = capon.getCapon(z, w, R, x, d, L, Yavg, M, Nx, Nz, Nb);
  if (e) { report("getCapon", e); return 1; }
  BuildR buildr;
  bool Rgpu = false, xgpu = false;
  e = buildr.getR(R2, x, d, L, Yavg, M, Nx, Nz, Nb, Rgpu, xgpu);
  if (e) { report("getR", e); return 1; }
  int N = 6, batch = 3;
  std::vector<cf> A((size_t)batch * N * N), b((size_t)batch * N), xs((size_t)batch * N);
  for (int s = 0; s < batch; ++s)
    for (int i = 0; i < N; ++i) {
      b[s * N + i] = cf(1.0f + i, 0.5f * s - i);
      for (int j = 0; j < N; ++j)
        A[(size_t)s * N * N + i * N + j] = cf((i == j ? 8.0f : 0.0f) + (float)std::sin(1.3 * (i + 2 * j + s)),
                                            (float)std::cos(0.7 * (3 * i - j + s)));
    }
  cf *xp = xs.data(), *Ap = A.data(), *bp = b.data();
  Solver solver;
  bool g0 = false, g1 = false, g2 = false;
  e = solver.solve(xp, Ap, bp, N, batch, g0, g1, g2);
  if (e) { report("solve", e); return 1; }
  put(dir + "/x.bin", xv);
  put(dir + "/z.bin", zv);
  put(dir + "/w.bin", wv);
  put(dir + "/R.bin", Rv);
  put(dir + "/Rg.bin", Rg);
  put(dir + "/A.bin", A);
  put(dir + "/b.bin", b);
  put(dir + "/xs.bin", xs);
  std::printf("ok\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 2 && !std::strcmp(argv[1], "errors")) return errors();
  if (argc >= 3 && !std::strcmp(argv[1], "run")) return run(argv[2]);
  std::fprintf(stderr, "usage: driver errors | driver run <dir>\n");
  return 2;
}
