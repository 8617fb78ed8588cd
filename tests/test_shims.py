"""The header-only C++ shims (include/Capon.h, BuildR.h, Solver.h) compiled into a driver
that calls them as the Cython code generated from RealTimeCapon/Capon.pxd, BuildR.pxd and
Solver.pxd does (one object, lvalue arguments bound to reference parameters), linked
against libcos_capon.so.  On the CPU: the reference's argument checks and their error
codes (no device call).  On the GPU: a small frame through all three classes against the
oracle and LAPACK."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "shim")
REF_COMPLEX = "/root/reference/framework/lib"     # the caller's Complex.h, where present


def _build(tmp_path, complex_dir=SHIM):
    exe = str(tmp_path / "driver")
    cmd = ["g++", "-std=c++11", "-O1", "-Wall", "-Werror", "-I", complex_dir, "-I", os.path.join(ROOT, "include"),
           os.path.join(SHIM, "driver.cpp"), "-o", exe, "-L", os.path.join(ROOT, "cos_amd"), "-lcos_capon",
           "-Wl,-rpath," + os.path.join(ROOT, "cos_amd")]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def test_shims_argument_errors(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = {l.split(" ", 2)[0]: (int(l.split(" ", 2)[1]), l.split(" ", 2)[2] if l.count(" ") >= 2 else "")
           for l in r.stdout.strip().splitlines()}
    expect = {"capon_L_gt_M": "L > M", "capon_2Y_ge_Nz": "2*Yavg >= Nz", "capon_Nb_gt_L": "Nb can not be larger than L",
              "buildr_Nz": "Nz <= Yavg", "solver_N": "N <= 72", "solver_batch": "batch >= 1"}
    for case, msg in expect.items():
        code, text = got[case]
        assert code == 1 and msg in text, (case, code, text)     # 1 = hipErrorInvalidValue


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_COMPLEX, "Complex.h")), reason="caller's Complex.h absent")
def test_shims_compile_against_the_callers_complex_h(tmp_path):
    _build(tmp_path, REF_COMPLEX)


@pytest.mark.gpu
def test_shims_run_a_frame(tmp_path):
    from oracle import capon_oracle as O
    print("[case] shim driver run", flush=True)
    exe = _build(tmp_path)
    r = subprocess.run([exe, "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
    ld = lambda n: np.fromfile(str(tmp_path / n), dtype=np.complex64)
    Nx, Nz, M, L, Y, d = 2, 9, 12, 5, 1, 0.2
    x = ld("x.bin").reshape(Nx, Nz, M)
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)
    z, w, R, Rg = ld("z.bin"), ld("w.bin").reshape(-1, L), ld("R.bin").reshape(ref["R"].shape), ld("Rg.bin")
    assert np.abs(z - ref["z"].ravel()).max() <= 2e-4 * np.abs(ref["z"]).max()
    assert np.abs(w - ref["w"].reshape(-1, L)).max() <= 2e-4 * np.abs(ref["w"]).max()
    assert np.abs(R - ref["R"]).max() <= 2e-6 * np.abs(ref["R"]).max() * np.sqrt(M)
    np.testing.assert_array_equal(Rg, R.ravel())                  # getR gives the same R
    N, batch = 6, 3
    A, b, xs = ld("A.bin").reshape(batch, N, N), ld("b.bin").reshape(batch, N), ld("xs.bin").reshape(batch, N)
    sol = np.linalg.solve(np.swapaxes(A, 1, 2).astype(complex), b.astype(complex)[..., None])[..., 0]
    np.testing.assert_allclose(xs, sol, rtol=0, atol=1e-4 * np.abs(sol).max())
