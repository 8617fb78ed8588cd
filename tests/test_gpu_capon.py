"""Parity of the HIP path (through the C ABI) with the oracle, on the GPU.

Tolerances are written per test.  The reference's own GPU test compared to its FP64
CPU path at 4 decimal places (framework/beamformer/capon/tests/testRTC.pyx:383-384);
here power is held to 1e-4 relative against FP64 on the 64-element, L = 32 workload.
"""
import sys

import numpy as np
import pytest

from oracle import c_capon
from oracle import capon_oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def uniform_cube(seed, Nx, Nz, M):
    # the reference test draws uniform complex data (testRTC.pyx:262-269), seeded here
    rng = np.random.default_rng(seed)
    return (rng.uniform(-1, 1, (Nx, Nz, M)) + 1j * rng.uniform(-1, 1, (Nx, Nz, M))).astype(np.complex64)


# ------------------------------------------------------------------ drop-in entry point

DROPIN_CASES = [
    # (M, L, Y, Nx, Nz, d) following testCaponGPU's grid (testRTC.pyx:262-313)
    (12, 2, 0, 2, 3, 0.2), (12, 3, 1, 2, 57, 0.2), (12, 6, 1, 10, 57, 0.2),
    (32, 5, 0, 10, 3, 0.2), (32, 8, 1, 2, 57, 0.2), (32, 16, 1, 10, 57, 0.2),
    (32, 32, 0, 3, 9, 0.2), (16, 1, 2, 2, 9, 0.2), (64, 32, 2, 4, 40, 0.01),
    (40, 20, 2, 3, 30, 0.05), (96, 48, 2, 1, 12, 0.01), (128, 64, 2, 1, 9, 0.01),
    (80, 37, 1, 2, 20, 0.05),   # 16 lanes x 3 columns with identity padding (37 < 48)
]


@pytest.mark.parametrize("M,L,Y,Nx,Nz,d", DROPIN_CASES)
def test_dropin_matches_oracle(M, L, Y, Nx, Nz, d):
    from cos_amd import getCaponCUDAPy
    say("dropin", M, L, Y, Nx, Nz, d)
    x = uniform_cube(100 + M + L, Nx, Nz, M)
    z, R, w, e = getCaponCUDAPy(x, d, L, Y, np.array([0]), False, False)
    assert e == 0
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)
    # amplitude and weights: relative to the image / weight scale
    assert np.max(np.abs(z - ref["z"])) <= 2e-4 * np.max(np.abs(ref["z"])) + 1e-6
    assert np.max(np.abs(w - ref["w"])) <= 2e-4 * np.max(np.abs(ref["w"]))
    # R: loaded, unnormalised, conj-stored, relative to its largest entry
    assert np.max(np.abs(R - ref["R"])) <= 2e-6 * np.max(np.abs(ref["R"])) * np.sqrt(M)


WIDE_CASES = [
    # (M, L, Y, Nx, Nz, d): subarrays past the fused kernel's 64 lanes, up to the reference
    # solver's 72 (RTC/Solver.cpp:34-42)
    (80, 65, 1, 2, 12, 0.2), (100, 70, 2, 2, 11, 0.05), (144, 72, 1, 3, 8, 0.01), (72, 72, 0, 2, 5, 0.2),
]


@pytest.mark.parametrize("M,L,Y,Nx,Nz,d", WIDE_CASES)
def test_dropin_wide_subarrays_match_oracle(M, L, Y, Nx, Nz, d):
    from cos_amd import getCaponCUDAPy, get_capon
    say("dropin wide", M, L, Y, Nx, Nz, d)
    x = uniform_cube(300 + M + L, Nx, Nz, M)
    z, R, w, e = getCaponCUDAPy(x, d, L, Y, np.array([0]), False, False)
    assert e == 0
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)
    assert np.max(np.abs(z - ref["z"])) <= 2e-4 * np.max(np.abs(ref["z"])) + 1e-6
    assert np.max(np.abs(w - ref["w"])) <= 2e-4 * np.max(np.abs(ref["w"]))
    assert np.max(np.abs(R - ref["R"])) <= 2e-6 * np.max(np.abs(ref["R"])) * np.sqrt(M)
    P = get_capon(x, d, L, Y)["P"].ravel()
    Pref = O.capon(x.astype(np.complex128), float(np.float32(d)), L, Y)["P"].ravel()
    assert rel(P, Pref).max() <= 1e-4


def test_dropin_wide_beamspace_and_limits():
    """L = 72 with Nb = 8 beams (the element-space S_L from the wide kernel, then B S_L Bᴴ),
    and L = 73 rejected with the solver's limit in the message."""
    from cos_amd import getCaponCUDAPy
    from cos_amd._lib import lib
    say("dropin wide beamspace", 80, 72, 8)
    x = uniform_cube(77, 2, 8, 80)
    z, R, w, e = getCaponCUDAPy(x, 0.2, 72, 1, np.zeros((8, 72)), False, False)
    assert e == 0
    ref = O.capon_beamspace_matrix(x.astype(np.complex128), float(np.float32(0.2)), 72, 1, 8)
    assert np.abs(z.ravel() - ref["z"]).max() <= 5e-4 * np.abs(ref["z"]).max()
    assert np.abs(w.reshape(-1, 72) - ref["w"]).max() <= 5e-4 * np.abs(ref["w"]).max()
    h = np.zeros(80 * 8 * 2, np.complex64)
    assert lib().cos_capon_get_capon_p(None, None, None, None, h.ctypes.data, 0.1, 73, 0, 80, 2, 4, 0) != 0
    assert b"L > 72" in lib().cos_capon_last_error()


@pytest.mark.parametrize("M,L", [(100, 100), (140, 128), (90, 65)])
def test_buildr_wide_subarrays(M, L):
    """BuildR::getR up to the reference's L = 128 (RTC/buildR_kernel.cu:260-264)."""
    from cos_amd import BuildR
    say("buildR wide", M, L)
    x = uniform_cube(9 + L, 2, 6, M)
    Nx, Nz, M = x.shape
    Y, d = 1, 0.1
    R = np.zeros((Nx, Nz - 2 * Y, L, L), np.complex64)
    assert BuildR().getR(R, x, d, L, Y, M, Nx, Nz, 0) == 0
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)["R"]
    assert np.abs(R - ref).max() <= 1e-5 * np.abs(ref).max()


def test_dropin_power_extension():
    from cos_amd import get_capon
    say("dropin power", 32, 16, 1)
    x = uniform_cube(7, 3, 20, 32)
    o = get_capon(x, 0.1, 16, 1, want_R=True)
    ref = O.capon(x.astype(np.complex128), float(np.float32(0.1)), 16, 1)
    assert np.max(rel(o["P"].ravel(), ref["P"])) < 1e-5


# ------------------------------------------------------------------ FP32 vs FP64 at the headline shape

def _cfg2_slice():
    from cos_amd import synth
    return synth.config2(Nx=8, Nz=2048, M=64, seed=2)


@pytest.fixture(scope="module")
def cfg2_slice():
    x = _cfg2_slice()
    ref = c_capon(x, 0.01, 32, 2)
    return x, ref


def _plan_run(x, d, L, Y, **kw):
    import torch
    from cos_amd import CaponPlan
    xt = torch.from_numpy(x).cuda()
    plan = CaponPlan()
    o = plan.run(xt, d, L, Y, **kw)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in o.items()}, plan


@pytest.mark.parametrize("cov", [False, "valu"])
def test_cfg2_slice_power_within_1e4_of_fp64(cfg2_slice, cov):
    """The default covariance (rank-two corrections on the matrix cores) and the recursion
    entirely on the VALU."""
    say("cfg2 slice fp32", cov, 8, 2048, 64, 32)
    x, ref = cfg2_slice
    o, _ = _plan_run(x, 0.01, 32, 2, cov_mfma=cov)
    e = rel(o["P"], ref["P"].ravel())
    print(f"P rel err max {e.max():.3e} p99.9 {np.quantile(e, 0.999):.3e}", flush=True)
    assert e.max() <= 1e-4
    we = np.linalg.norm(o["w"] - ref["w"].reshape(-1, 32), axis=1) / np.linalg.norm(ref["w"].reshape(-1, 32), axis=1)
    print(f"w rel err max {we.max():.3e}", flush=True)
    assert we.max() <= 1e-3
    ze = np.abs(o["z"] - ref["z"].ravel()) / np.abs(ref["z"].ravel()).max()
    assert ze.max() <= 1e-4


def test_cfg2_slice_fp64_kernel(cfg2_slice):
    say("cfg2 slice fp64", 8, 2048, 64, 32)
    x, ref = cfg2_slice
    o, _ = _plan_run(x, 0.01, 32, 2, precision="fp64")
    # FP64 arithmetic, float32 outputs: the bound is the float32 rounding of the result
    assert rel(o["P"], ref["P"].ravel()).max() <= 1.2e-7
    np.testing.assert_allclose(o["w"], ref["w"].reshape(-1, 32), rtol=0, atol=2e-7 * np.abs(ref["w"]).max())


# ------------------------------------------------------------------ iterative refinement

def test_cfg2_slice_refined_power(cfg2_slice):
    """One refinement step (residual from the snapshots) on the 64-element, L = 32 slice:
    power within 1e-4 of FP64 and no worse than the unrefined FP32 result."""
    say("cfg2 slice fp32 refine", 8, 2048, 64, 32)
    x, ref = cfg2_slice
    o0, _ = _plan_run(x, 0.01, 32, 2)
    o1, _ = _plan_run(x, 0.01, 32, 2, refine=True)
    e0, e1 = rel(o0["P"], ref["P"].ravel()), rel(o1["P"], ref["P"].ravel())
    w64 = ref["w"].reshape(-1, 32)
    we = lambda w: np.linalg.norm(w - w64, axis=1) / np.linalg.norm(w64, axis=1)
    print(f"P err max {e0.max():.3e} -> {e1.max():.3e}, mean {e0.mean():.3e} -> {e1.mean():.3e}; "
          f"w err max {we(o0['w']).max():.3e} -> {we(o1['w']).max():.3e}", flush=True)
    assert e1.max() <= 1e-4
    assert e1.max() <= max(e0.max(), 2e-6)
    assert we(o1["w"]).max() <= 1e-3
    ze = np.abs(o1["z"] - ref["z"].ravel()) / np.abs(ref["z"].ravel()).max()
    assert ze.max() <= 1e-4


@pytest.mark.parametrize("mode", ["fused", "rank", "valu"])
@pytest.mark.parametrize("M,L,Y,Nx,Nz", [(64, 32, 2, 2, 40), (48, 24, 1, 3, 37), (40, 20, 0, 2, 30), (80, 32, 2, 2, 44),
                                       (33, 17, 2, 1, 21), (64, 32, 2, 3, 43), (120, 32, 2, 2, 40), (96, 24, 1, 2, 34)])
def test_fused_mfma_covariance_matches_oracle(M, L, Y, Nx, Nz, mode):
    """Covariance work on the matrix cores inside the fused kernel ("fused": per-snapshot
    Grams from an FP16 hi/lo split of power-of-two scaled rows; "rank": the recursion's
    rank-two corrections per window from FP16 hi/lo-split products of scaled rows, rows
    padded past 128 elements (M = 120, 96) through the kernel that stages them; "valu":
    none, the recursion on the VALU): P, w and z against the FP64 oracle, on shapes whose
    waves take the matrix-core path (four consecutive rows) and, where the row count is not
    a multiple of four, the recursion at the beam ends."""
    say("fused mfma covariance", mode, M, L, Y, Nx, Nz)
    x = uniform_cube(M * 7 + L + Y, Nx, Nz, M)
    o, _ = _plan_run(x, 0.05, L, Y, cov_mfma=mode)
    ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
    eP = rel(o["P"], ref["P"])
    we = np.linalg.norm(o["w"] - ref["w"], axis=1) / np.linalg.norm(ref["w"], axis=1)
    ez = np.abs(o["z"] - ref["z"]).max() / np.abs(ref["z"]).max()
    print(f"P {eP.max():.2e} w {we.max():.2e} z {ez:.2e}", flush=True)
    assert eP.max() <= 1e-5 and we.max() <= 1e-5 and ez <= 1e-5


@pytest.mark.parametrize("mode", ["fused", "rank"])
def test_fused_mfma_covariance_headline_slice(cfg2_slice, mode):
    """The 64-element, L = 32 slice (speckle and 30 dB scatterers spanning 2^5 in amplitude
    within a wave's rows) through the matrix-core covariance: power within 1e-5 of FP64."""
    say("fused mfma covariance cfg2 slice", mode, 8, 2048, 64, 32)
    x, ref = cfg2_slice
    o, _ = _plan_run(x, 0.01, 32, 2, cov_mfma=mode)
    e = rel(o["P"], ref["P"].ravel())
    print(f"P rel err max {e.max():.3e}", flush=True)
    assert e.max() <= 1e-5


@pytest.mark.parametrize("e", [-40, 50])
@pytest.mark.parametrize("M,L,Y", [(64, 32, 2), (120, 32, 2), (128, 64, 2), (96, 40, 2)])
def test_split_recursion_is_scale_free(M, L, Y, e):
    """The FP16 hi/lo split of the rank-two products works on rows normalised by a power of
    two (the largest component into [2^7, 2^8)), so data multiplied by 2^e must give P times
    2^(2e), z times 2^e and the same w, bit for bit: the split sees the same operands and
    every FP32 step on the scaled data is exact.  Shapes: the 16-lane kernel (M = 64), its
    wide-row instantiation (M = 120), the 64-lane kernel (L = 64, its scale folded into the
    first row and the trace), three columns per lane (L = 40: rows and columns 0..31 on the
    matrix cores, the third slot on the VALU); Nz - 2Y a multiple of four, so every wave
    takes the path.
    (2^-40 and 2^50 keep P a normal float: its scale-back is then exact.)"""
    say("split recursion scale", M, L, Y, e)
    x = uniform_cube(M + 5 * L + Y, 2, 32 + 2 * Y, M)
    xs = (x * np.float32(2.0 ** e)).astype(np.complex64)
    o0, _ = _plan_run(x, 0.05, L, Y)
    o1, _ = _plan_run(xs, 0.05, L, Y)
    assert np.array_equal(o1["P"], (o0["P"].astype(np.float64) * 2.0 ** (2 * e)).astype(np.float32))
    assert np.array_equal(o1["z"], (o0["z"].astype(np.complex128) * 2.0 ** e).astype(np.complex64))
    assert np.array_equal(o1["w"], o0["w"])
    ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
    assert rel(o0["P"], ref["P"]).max() <= 1e-5


@pytest.mark.parametrize("e", [-40, 50])
def test_split_recursion_refined_is_scale_free(e):
    """The refinement kernel on the split recursion forms its residual from the unscaled
    rows and the scaled solve: data times 2^e must still give P times 2^(2e), z times 2^e
    and the same w, bit for bit, and the refined power must beat the unrefined one."""
    say("split recursion refined scale", 64, 32, 2, e)
    x = uniform_cube(417, 2, 36, 64)
    xs = (x * np.float32(2.0 ** e)).astype(np.complex64)
    o0, _ = _plan_run(x, 0.05, 32, 2, refine=True)
    o1, _ = _plan_run(xs, 0.05, 32, 2, refine=True)
    assert np.array_equal(o1["P"], (o0["P"].astype(np.float64) * 2.0 ** (2 * e)).astype(np.float32))
    assert np.array_equal(o1["z"], (o0["z"].astype(np.complex128) * 2.0 ** e).astype(np.complex64))
    assert np.array_equal(o1["w"], o0["w"])
    ref = O.capon(x.astype(np.complex128), 0.05, 32, 2)
    u, _ = _plan_run(x, 0.05, 32, 2)
    assert rel(o0["P"], ref["P"]).max() <= max(rel(u["P"], ref["P"]).max(), 1e-7)


def test_rank_corrections_covariance_output():
    """The rank-two corrections on the matrix cores with the materialised R (the
    reference layout) against the oracle's."""
    import torch
    from cos_amd import CaponPlan
    say("rank covariance R + refine", 64, 32, 2)
    x = uniform_cube(91, 2, 30, 64)
    xt = torch.from_numpy(x).cuda()
    o = CaponPlan().run(xt, 0.05, 32, 2, cov_mfma="rank", want=("P", "z", "w", "R"))
    torch.cuda.synchronize()
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(0.05)), 32, 2)
    R = o["R"].cpu().numpy().reshape(ref["R"].shape)
    assert np.abs(R - ref["R"]).max() <= 2e-6 * np.abs(ref["R"]).max() * 8
    assert rel(o["P"].cpu().numpy(), ref["P"].ravel()).max() <= 1e-5


def test_refinement_on_ill_conditioned_covariance():
    """d = 1e-4 (condition bound 1 + L/d = 3.2e5, below the fallback threshold): the FP32
    solve loses digits; one refinement step must bring power and weights closer to FP64."""
    from cos_amd import synth
    say("refine ill-conditioned", 4, 64, 64, 32)
    x = synth.config2(Nx=4, Nz=64, M=64, seed=12)
    ref = c_capon(x, 1e-4, 32, 2)
    o0, _ = _plan_run(x, 1e-4, 32, 2)
    o1, _ = _plan_run(x, 1e-4, 32, 2, refine=True)
    e0, e1 = rel(o0["P"], ref["P"].ravel()), rel(o1["P"], ref["P"].ravel())
    w64 = ref["w"].reshape(-1, 32)
    we = lambda w: np.linalg.norm(w - w64, axis=1) / np.linalg.norm(w64, axis=1)
    print(f"P err max {e0.max():.3e} -> {e1.max():.3e}; w err max {we(o0['w']).max():.3e} -> "
          f"{we(o1['w']).max():.3e}", flush=True)
    assert e1.max() <= 1e-4
    assert e1.mean() <= e0.mean()
    assert np.median(we(o1["w"])) <= np.median(we(o0["w"]))


@pytest.mark.parametrize("M,L,Y", [(24, 12, 1), (48, 24, 2), (80, 40, 1), (128, 64, 2), (128, 64, 1), (16, 1, 2),
                                   (32, 32, 1), (20, 20, 0)])
def test_refined_shapes_match_oracle(M, L, Y):
    """The refined kernel in the default lane shape of each L (16 lanes x 1, 2 or 3 columns,
    64 lanes) against the FP64 oracle: P within 1e-4, w within 1e-3 of its largest entry,
    z within 1e-4.  At L = 64 (blocks of four 64-lane waves) a beam of 14 rows holds 10
    pixels with Y = 2, so some blocks straddle two beams and stage their rows per wave, and
    12 with Y = 1, so every block shares its window rows."""
    say("refine shape", M, L, Y)
    x = uniform_cube(3 * M + L, 3, 14, M)
    o, _ = _plan_run(x, 0.05, L, Y, refine=True)
    ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
    eP = rel(o["P"], ref["P"].ravel())
    ez = np.abs(o["z"] - ref["z"].ravel()) / np.abs(ref["z"]).max()
    print(f"P err {eP.max():.2e} z err {ez.max():.2e}", flush=True)
    assert eP.max() <= 1e-4
    assert ez.max() <= 1e-4
    np.testing.assert_allclose(o["w"], ref["w"].reshape(-1, L), atol=1e-3 * np.abs(ref["w"]).max())


def test_refined_unloaded_power_and_power_amplitude():
    say("refine options", 48, 24, 2)
    x = uniform_cube(77, 3, 12, 48)
    o, _ = _plan_run(x, 0.05, 24, 2, refine=True, z_mode="power", p_mode="unloaded")
    ref = O.capon_options(x.astype(np.complex128), 0.05, 24, 2, fb=False, z_mode=1, p_mode=1)
    assert rel(o["P"], ref["P"]).max() <= 1e-4
    assert (np.abs(o["z"] - ref["z"]) / np.abs(ref["z"]).max()).max() <= 1e-4


def test_refined_rows_tables_with_fallback():
    """Per-row apertures (L 8..32) with 7 rows at d = 1e-5: the refined FP32 runs, then the
    FP64 pass over the flagged pixels."""
    import torch
    from cos_amd import synth
    say("refine rows+fallback", 4, 96, 64)
    Nx, Nz, Mmax, Y = 4, 96, 64, 2
    Mz = (2 * np.round((16 + (Mmax - 16) * np.arange(Nz) / (Nz - 1)) / 2)).astype(np.int32)
    Lz = (Mz // 2).astype(np.int32)
    off = ((Mmax - Mz) // 2).astype(np.int32)
    dz = np.full(Nz, 0.01, np.float32)
    dz[10:14] = 1e-5
    dz[60:63] = 1e-5
    x = synth.speckle_cube(Nx, Nz, Mmax, seed=31)
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    o, plan = _plan_run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", refine=True,
                        want=("P", "z", "w", "flags"))
    ref = O.capon_rows(x.astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y)
    e = rel(o["P"], ref["P"])
    print(f"refined rows P rel err max {e.max():.3e}", flush=True)
    assert e.max() <= 1e-4
    flagged = np.nonzero(o["flags"])[0]
    assert plan.fallback_count() == len(flagged) > 0
    assert rel(o["P"][flagged], ref["P"][flagged]).max() <= 1.2e-7


# ------------------------------------------------------------------ matrix-core covariance

@pytest.mark.parametrize("M,L,Y", [(64, 32, 2), (24, 12, 1), (40, 17, 0), (16, 1, 2), (80, 40, 1), (128, 64, 2),
                                   (70, 33, 2)])
def test_mfma_covariance_matches_oracle(M, L, Y):
    """The covariance on the matrix cores (BF16 hi/lo split, FP32 accumulation) in the
    reference R layout (loaded, unnormalised, conj-stored) against the FP64 oracle: within
    2e-5 of the largest entry (three BF16 products keep 16 significant bits of each
    operand; measured 3-4e-6); the solve from it gives P within 2e-4 of FP64."""
    say("mfma covariance", M, L, Y)
    x = uniform_cube(11 * M + L, 3, 10, M)
    o, _ = _plan_run(x, 0.05, L, Y, cov_mfma=True, want=("P", "z", "w", "R"))
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(0.05)), L, Y)
    Rref = ref["R"].reshape(o["R"].shape)
    eR = np.abs(o["R"] - Rref).max() / np.abs(Rref).max()
    refP = O.capon(x.astype(np.complex128), 0.05, L, Y)
    eP = rel(o["P"], refP["P"].ravel())
    print(f"R err {eR:.2e} (of max), P err max {eP.max():.2e}", flush=True)
    assert eR <= 2e-5
    assert eP.max() <= 2e-4


def test_mfma_covariance_cfg2_slice(cfg2_slice):
    """The 64-element, L = 32 slice through the matrix-core covariance: power within 1e-4
    of FP64 (measured max 5e-5, mean 5e-6)."""
    say("cfg2 slice mfma covariance", 8, 2048, 64, 32)
    x, ref = cfg2_slice
    o, _ = _plan_run(x, 0.01, 32, 2, cov_mfma=True)
    e = rel(o["P"], ref["P"].ravel())
    print(f"mfma-covariance P err max {e.max():.3e} mean {e.mean():.3e}", flush=True)
    assert e.max() <= 1e-4


def test_mfma_covariance_aperture_rows():
    """The expanding-aperture geometry (L 8..64 by row, 5% of rows at d = 1e-5) through the
    matrix-core covariance, one launch per row run with one or four tiles, then the FP32
    solves and the FP64 pass over the flagged pixels: power within 1e-4 of FP64."""
    import torch
    from cos_amd import synth
    say("mfma covariance aperture", 2, 1024, 128)
    Nx, Nz, Y = 2, 1024, 2
    Mz, Lz, off, dz = synth.config3_geometry(Nz, seed=3, Mmax=128)
    x = synth.speckle_cube(Nx, Nz, 128, seed=3)
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    o, plan = _plan_run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", cov_mfma=True, want=("P", "w", "flags"))
    Nzp = Nz - 2 * Y
    sel = np.arange(0, Nx * Nzp, 5)
    ref = O.capon_rows(x.astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y,
                       pixels=(sel // Nzp, sel % Nzp))
    e = rel(o["P"][sel], ref["P"])
    print(f"mfma aperture P rel err max {e.max():.3e}; fallback {plan.fallback_count()}", flush=True)
    assert e.max() <= 1e-4
    assert plan.fallback_count() > 0


# ------------------------------------------------------------------ lane-shape autotuner

SHAPE_CAP = (16, 32, 64, 48)   # lane shapes 1..4: the largest L each holds


@pytest.mark.parametrize("M,L,Y", [(24, 12, 2), (64, 32, 2), (96, 40, 2), (96, 48, 2), (80, 36, 4), (56, 33, 0),
                                   (48, 48, 1), (40, 34, 2), (32, 32, 2)])
def test_every_lane_shape_matches_oracle(M, L, Y):
    """Each lane shape that holds L, forced through the tuning table, against the FP64
    oracle (windows of 1, 5 and 9 rows: the three fixed-window kernels); the table is
    reset afterwards."""
    import torch
    from cos_amd import CaponPlan
    x = uniform_cube(5 * M + L, 3, 20, M)
    xt = torch.from_numpy(x).cuda()
    ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
    try:
        for sh in (1, 2, 3, 4):
            if SHAPE_CAP[sh - 1] < L:
                continue
            say("forced shape", M, L, Y, sh)
            CaponPlan.set_tuning(xt, 0.05, L, Y, {L: sh})
            o = CaponPlan().run(xt, 0.05, L, Y, precision="fp32+fallback")
            torch.cuda.synchronize()
            assert rel(o["P"].cpu().numpy(), ref["P"].ravel()).max() <= 1e-4
            np.testing.assert_allclose(o["w"].cpu().numpy(), ref["w"].reshape(-1, L),
                                       atol=1e-3 * np.abs(ref["w"]).max())
            np.testing.assert_allclose(o["z"].cpu().numpy(), ref["z"].ravel(), atol=1e-4 * np.abs(ref["z"]).max())
    finally:
        CaponPlan.set_tuning(xt, 0.05, L, Y, {})


@pytest.mark.parametrize("M,L,Y", [(80, 40, 1), (96, 48, 4), (72, 36, 2)])
def test_three_column_shape_refined(M, L, Y):
    """Refinement in lane shape 4 (the fixed five-row and runtime-window kernels), forced
    through the tuning table, against the FP64 oracle and against the same frame refined
    by the 64-lane shape: P within 1e-4 (and at least as close on average as unrefined),
    w within 1e-3 of its largest entry, z within 1e-4."""
    import torch
    from cos_amd import CaponPlan
    say("shape 4 refined", M, L, Y)
    x = uniform_cube(2 * M + L + Y, 3, 14, M)
    xt = torch.from_numpy(x).cuda()
    ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
    out = {}
    try:
        for sh in (3, 4):
            for refine in (False, True):
                CaponPlan.set_tuning(xt, 0.05, L, Y, {L: sh}, refine=refine)
                out[sh, refine], _ = _plan_run(x, 0.05, L, Y, refine=refine)
    finally:
        for refine in (False, True):
            CaponPlan.set_tuning(xt, 0.05, L, Y, {}, refine=refine)
    e = {k: rel(o["P"], ref["P"].ravel()) for k, o in out.items()}
    print({k: f"{v.max():.2e}/{v.mean():.2e}" for k, v in e.items()}, flush=True)
    o = out[4, True]
    assert e[4, True].max() <= 1e-4
    assert e[4, True].mean() <= e[4, False].mean()
    assert e[4, True].mean() <= 2 * e[3, True].mean() + 1e-9
    assert np.abs(o["z"] - ref["z"].ravel()).max() <= 1e-4 * np.abs(ref["z"]).max()
    np.testing.assert_allclose(o["w"], ref["w"].reshape(-1, L), atol=1e-3 * np.abs(ref["w"]).max())


@pytest.mark.parametrize("case", ["options", "steering", "mfma_covariance"])
def test_three_column_shape_runtime_window_kernel(case):
    """Lane shape 4 (16 lanes x 3 columns) in its runtime-window kernel: the reference
    options (forward-backward averaging, power amplitude, unloaded power), steering
    vectors, and the solve from a stored covariance, each against the FP64 oracle."""
    import torch
    from cos_amd import CaponPlan, synth
    M, L, Y = 80, 40, 1
    say("shape 4", case, M, L, Y)
    x = uniform_cube(M + 3 * L, 3, 12, M)
    xt = torch.from_numpy(x).cuda()
    opts = dict(fb_avg=True, z_mode="power", p_mode="unloaded") if case == "options" else {}
    try:
        CaponPlan.set_tuning(xt, 0.05, L, Y, {L: 4}, **opts)
        if case == "options":
            o, _ = _plan_run(x, 0.05, L, Y, **opts)
            ref = O.capon_options(x.astype(np.complex128), 0.05, L, Y, fb=True, z_mode=1, p_mode=1)
            assert rel(o["P"], ref["P"]).max() <= 1e-4
            assert np.abs(o["z"] - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max()
            np.testing.assert_allclose(o["w"].reshape(-1, L), ref["w"], atol=1e-3 * np.abs(ref["w"]).max())
        elif case == "steering":
            sines = synth.steering_sines(8).astype(np.float32)
            o, _ = _plan_run(x, 0.05, L, Y, steer_sin=torch.from_numpy(sines).cuda())
            ref = O.capon_steered(x.astype(np.complex128), 0.05, L, Y, sines.astype(np.float64))
            assert rel(o["P"].reshape(-1, len(sines)), ref["P"]).max() <= 1e-4
            w = o["w"].reshape(-1, len(sines), L)
            assert np.abs(w - ref["w"]).max() <= 1e-3 * np.abs(ref["w"]).max()
        else:
            o, _ = _plan_run(x, 0.05, L, Y, cov_mfma=True)
            ref = O.capon(x.astype(np.complex128), 0.05, L, Y)
            assert rel(o["P"], ref["P"].ravel()).max() <= 1e-4
    finally:
        CaponPlan.set_tuning(xt, 0.05, L, Y, {}, **opts)


def test_autotune_uniform_frame():
    """The autotuner on a frame of one subarray length: the three shapes that hold L = 24
    are timed, the fastest is kept, and the tuned frame matches FP64."""
    import torch
    from cos_amd import CaponPlan
    say("autotune uniform", 48, 24)
    x = uniform_cube(91, 4, 160, 48)
    xt = torch.from_numpy(x).cuda()
    plan = CaponPlan()
    try:
        tuned = plan.autotune(xt, 0.05, 24, 2)
        assert list(tuned["shapes"]) == [24] and sorted(tuned["ms"][24]) == [2, 3, 4]
        assert tuned["ms"][24][tuned["shapes"][24]] == min(tuned["ms"][24].values())
        o = plan.run(xt, 0.05, 24, 2, precision="fp32+fallback")
        torch.cuda.synchronize()
        ref = O.capon(x.astype(np.complex128), 0.05, 24, 2)
        assert rel(o["P"].cpu().numpy(), ref["P"].ravel()).max() <= 1e-4
    finally:
        CaponPlan.set_tuning(xt, 0.05, 24, 2, {})


def test_autotune_aperture_rows():
    """The autotuner over an expanding-aperture frame (L 8..64): it times every shape that
    holds each length, installs the fastest, and the tuned frame still matches FP64; a
    captured graph is re-captured after the table changes."""
    import torch
    from cos_amd import CaponPlan, synth
    say("autotune aperture", 2, 1024, 128)
    Nx, Nz, Y = 2, 1024, 2
    Mz, Lz, off, dz = synth.config3_geometry(Nz, seed=3, Mmax=128)
    x = synth.speckle_cube(Nx, Nz, 128, seed=3)
    xt = torch.from_numpy(x).cuda()
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    plan = CaponPlan()
    try:
        out = plan.run(xt, 0.0, 0, Y, rows=rows, precision="fp32+fallback", graph=True)
        tuned = plan.autotune(xt, 0.0, 0, Y, rows=rows)
        Ls = sorted(set(Lz[Y:Nz - Y].tolist()))
        assert sorted(tuned["shapes"]) == Ls
        for l, sh in tuned["shapes"].items():
            valid = [s for s in (1, 2, 3, 4) if SHAPE_CAP[s - 1] >= l]
            assert sorted(tuned["ms"][l]) == valid
            assert tuned["ms"][l][sh] == min(tuned["ms"][l].values())
        print("tuned shapes", tuned["shapes"], flush=True)
        o = plan.run(xt, 0.0, 0, Y, rows=rows, precision="fp32+fallback", graph=True, out=out)
        torch.cuda.synchronize()
        Nzp = Nz - 2 * Y
        sel = np.arange(0, Nx * Nzp, 7)
        ref = O.capon_rows(x.astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y,
                           pixels=(sel // Nzp, sel % Nzp))
        assert rel(o["P"].cpu().numpy()[sel], ref["P"]).max() <= 1e-4
    finally:
        CaponPlan.set_tuning(xt, 0.0, 0, Y, {}, rows=rows)


# ------------------------------------------------------------------ per-row apertures and fallback

def test_fallback_list_past_the_expected_count():
    """More flagged pixels than the host expects (FP32 breakdowns on rank-deficient windows,
    none flagged by the condition bound): the list entries past the expected count go to
    the persistent FP64 grid, and every flagged pixel matches the all-FP64 frame.  The list
    kernel forms the covariance by the FP64 recursion, the whole-frame kernel by direct
    sums, so the two agree to the window's conditioning: with two subarrays S has rank 2,
    and the loading d = 1e-8 gives cond(S_L) ~ L/(2d) = 1.6e9 (FP32 pivots drown in
    rounding, FP64 ones keep ~7 digits)."""
    M, L, Y, d = 33, 32, 0, 1e-8   # two subarrays: S_L has rank 2 plus a loading of 1e-8
    say("fallback tail", 8, 2048, M, L, Y, d)
    x = uniform_cube(33, 8, 2048, M)
    o, plan = _plan_run(x, d, L, Y, precision="fp32+fallback", cond_thresh=1e30, want=("P", "z", "w", "flags"))
    n = plan.fallback_count()
    flagged = np.nonzero(o["flags"])[0]
    print(f"flagged {n} of {o['P'].size}", flush=True)
    assert n == len(flagged)
    assert n > 4096 + 1000          # past the head launch (host count 0 + 4096)
    assert np.all((o["flags"][flagged] & 2) != 0)   # breakdowns, not the bound
    ref, _ = _plan_run(x, d, L, Y, precision="fp64", want=("P", "z", "w"))
    assert np.all(np.isfinite(o["P"][flagged]))
    np.testing.assert_allclose(o["P"][flagged], ref["P"][flagged], rtol=1e-5, atol=0)
    np.testing.assert_allclose(o["w"][flagged], ref["w"][flagged], rtol=0, atol=1e-5 * np.abs(ref["w"][flagged]).max())
    np.testing.assert_allclose(o["z"][flagged], ref["z"][flagged], rtol=0, atol=1e-5 * np.abs(ref["z"][flagged]).max())


def test_rows_tables_and_fp64_fallback():
    import torch
    from cos_amd import synth
    say("rows+fallback", 4, 96, 64)
    Nx, Nz, Mmax, Y = 4, 96, 64, 2
    Mz = (2 * np.round((16 + (Mmax - 16) * np.arange(Nz) / (Nz - 1)) / 2)).astype(np.int32)
    Lz = (Mz // 2).astype(np.int32)
    off = ((Mmax - Mz) // 2).astype(np.int32)
    dz = np.full(Nz, 0.01, np.float32)
    dz[10:14] = 1e-5
    dz[60:63] = 1e-5
    x = synth.speckle_cube(Nx, Nz, Mmax, seed=31)
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    o, plan = _plan_run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", want=("P", "z", "w", "flags"))
    ref = O.capon_rows(x.astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y)
    e = rel(o["P"], ref["P"])
    print(f"rows P rel err max {e.max():.3e}", flush=True)
    assert e.max() <= 1e-4
    flagged = np.nonzero(o["flags"])[0]
    c = np.tile(np.arange(Nz - 2 * Y) + Y, Nx)
    expect = np.nonzero((1 + Lz[c] / dz[c].astype(np.float64)) > 1e6)[0]
    assert set(expect.tolist()) <= set(flagged.tolist())
    assert plan.fallback_count() == len(flagged)
    # flagged pixels were recomputed in FP64 (outputs stored as float32)
    assert rel(o["P"][flagged], ref["P"][flagged]).max() <= 1.2e-7
    for n in range(0, len(ref["w"]), 37):
        Ln = len(ref["w"][n])
        np.testing.assert_allclose(o["w"][n, :Ln], ref["w"][n], atol=2e-3 * np.abs(ref["w"][n]).max())


def test_expanding_aperture_full_depth():
    """The expanding-aperture geometry over its full depth (M 16..128, L 8..64, 5% of rows
    at d = 1e-5) on two beams: the frame is split into runs by subarray-length bucket,
    flagged pixels with L up to 64 go through the FP64 pass.  Every 5th pixel against the
    FP64 oracle; power within 1e-4, FP64-recomputed pixels within float32 rounding."""
    import torch
    from cos_amd import synth
    say("expanding aperture full depth", 2, 4096, 128)
    Nx, Nz, Y = 2, 4096, 2
    Mz, Lz, off, dz = synth.config3_geometry(Nz, seed=3, Mmax=128)
    x = synth.speckle_cube(Nx, Nz, 128, seed=3)
    m = np.arange(128)
    x *= ((m[None, :] >= off[:, None]) & (m[None, :] < (off + Mz)[:, None]))[None]
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    o, plan = _plan_run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", want=("P", "w", "flags"))
    Nzp = Nz - 2 * Y
    sel = np.arange(0, Nx * Nzp, 5)
    ref = O.capon_rows(x.astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y,
                       pixels=(sel // Nzp, sel % Nzp))
    e = rel(o["P"][sel], ref["P"])
    print(f"aperture P rel err max {e.max():.3e}; fallback {plan.fallback_count()}", flush=True)
    assert e.max() <= 1e-4
    c = np.tile(np.arange(Nzp) + Y, Nx)
    ill = (1 + Lz[c] / dz[c].astype(np.float64)) > 1e6
    flagged = np.nonzero(o["flags"])[0]
    assert set(np.nonzero(ill)[0].tolist()) <= set(flagged.tolist())
    assert (Lz[c[flagged]] > 32).any(), "the FP64 pass must see subarrays longer than 32"
    fsel = np.isin(sel, flagged)
    assert fsel.any()
    assert rel(o["P"][sel[fsel]], ref["P"][fsel]).max() <= 1.2e-7
    for n in range(0, len(sel), 97):
        Ln = len(ref["w"][n])
        np.testing.assert_allclose(o["w"][sel[n], :Ln], ref["w"][n], atol=2e-3 * np.abs(ref["w"][n]).max())
        assert np.all(o["w"][sel[n], Ln:] == 0)


# ------------------------------------------------------------------ steering vectors

def test_steering_matches_oracle():
    import torch
    from cos_amd import synth
    say("steering", 1, 64, 32, 16)
    x = synth.config1(N=64, M=32, seed=1)
    sines = synth.steering_sines(16).astype(np.float32)
    st = torch.from_numpy(sines).cuda()
    o, _ = _plan_run(x, 0.01, 16, 2, steer_sin=st)
    ref = O.capon_steered(x.astype(np.complex128), 0.01, 16, 2, sines.astype(np.float64))
    S = len(sines)
    P = o["P"].reshape(-1, S)
    assert rel(P, ref["P"]).max() <= 1e-4
    w = o["w"].reshape(-1, S, 16)
    assert np.abs(w - ref["w"]).max() <= 1e-3 * np.abs(ref["w"]).max()
    z = o["z"].reshape(-1, S)
    assert np.abs(z - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max()


def test_steering_full_bins_refined_within_1e4():
    """The whole far-field workload (1020 range bins x 128 angles, sources 20 dB above the
    noise): one refinement step with the steering vector as right-hand side holds every
    power within 1e-4 of FP64 (the unrefined FP32 solve reaches ~1.3e-4 on a few bins)."""
    import torch
    from cos_amd import synth
    say("steering full, refined", 1, 1024, 32, 16, 128)
    x = synth.config1(N=1024, M=32, seed=1)
    sines = synth.steering_sines(128).astype(np.float32)
    st = torch.from_numpy(sines).cuda()
    ref = O.capon_steered(x.astype(np.complex128), 0.01, 16, 2, sines.astype(np.float64))
    o0, _ = _plan_run(x, 0.01, 16, 2, steer_sin=st)
    o1, _ = _plan_run(x, 0.01, 16, 2, steer_sin=st, refine=True)
    S = len(sines)
    e0 = rel(o0["P"].reshape(-1, S), ref["P"])
    e1 = rel(o1["P"].reshape(-1, S), ref["P"])
    w1 = o1["w"].reshape(-1, S, 16)
    we = np.linalg.norm(w1 - ref["w"], axis=-1) / np.linalg.norm(ref["w"], axis=-1)
    print(f"steered P err max {e0.max():.3e} -> refined {e1.max():.3e}; refined w err max {we.max():.3e}",
          flush=True)
    assert e1.max() <= 1e-4
    assert e1.max() <= e0.max()
    assert we.max() <= 1e-3
    z1 = o1["z"].reshape(-1, S)
    assert np.abs(z1 - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max()


@pytest.mark.parametrize("L", [16, 40])
def test_steering_with_fp64_fallback(L):
    """Steered frames on the FP64 paths: per-row loading 1e-6 on every third range bin
    (condition bound 1 + L/d past 1e6) sends those bins to the FP64 kernel, which solves
    every angle of the flagged pixel; all powers, weights and amplitudes match the oracle
    (refined FP32 elsewhere), and the flagged bins match an all-FP64 steered run bit for
    bit (16 lanes and 64 lanes in FP64)."""
    import torch
    from cos_amd import synth
    M, Y = 2 * L, 2
    say("steering + fallback", 1, 256, M, L, 12)
    x = synth.config1(N=256, M=M, seed=1)
    sines = synth.steering_sines(12).astype(np.float32)
    st = torch.from_numpy(sines).cuda()
    Nz = x.shape[1]
    d_rows = np.full(Nz, 0.01, np.float32)
    d_rows[::3] = 1e-6
    rows = {"L": torch.full((Nz,), L, dtype=torch.int32).cuda(), "M": torch.full((Nz,), M, dtype=torch.int32).cuda(),
            "off": torch.zeros(Nz, dtype=torch.int32).cuda(), "d": torch.from_numpy(d_rows).cuda()}
    # the FP32 pixels take one refinement step (unrefined steered FP32 reaches ~1.3e-4 next
    # to the 20 dB sources); the flagged ones go to the FP64 kernel
    o, plan = _plan_run(x, 0.0, 0, Y, rows=rows, steer_sin=st, precision="fp32+fallback", refine=True,
                        want=("P", "z", "w", "flags"))
    o64, _ = _plan_run(x, 0.0, 0, Y, rows=rows, steer_sin=st, precision="fp64")
    S = len(sines)
    flagged = np.nonzero(o["flags"])[0]
    c = np.arange(Nz - 2 * Y) + Y
    assert plan.fallback_count() == len(flagged) == int((d_rows[c] < 1e-3).sum())
    P, P64 = o["P"].reshape(-1, S), o64["P"].reshape(-1, S)
    np.testing.assert_array_equal(P[flagged], P64[flagged])     # the same FP64 kernel
    np.testing.assert_array_equal(o["w"].reshape(-1, S, L)[flagged], o64["w"].reshape(-1, S, L)[flagged])
    for dval in (0.01, 1e-6):
        sel = np.nonzero(d_rows[c] == np.float32(dval))[0]
        ref = O.capon_steered(x.astype(np.complex128), float(np.float32(dval)), L, Y, sines.astype(np.float64),
                              pixels=(np.zeros(len(sel), int), sel))
        tolP = 1e-4 if dval > 1e-3 else 1e-6
        assert rel(P[sel], ref["P"]).max() <= tolP, dval
        assert rel(P64[sel], ref["P"]).max() <= 1e-6, dval
        w = o["w"].reshape(-1, S, L)[sel]
        assert np.abs(w - ref["w"]).max() <= 1e-3 * np.abs(ref["w"]).max(), dval
        z = o["z"].reshape(-1, S)[sel]
        assert np.abs(z - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max(), dval


# ------------------------------------------------------------------ beamspace drop-in (Nb > 0)

@pytest.mark.parametrize("M,L,Nb", [(12, 6, 2), (32, 16, 8), (32, 16, 9), (64, 32, 5)])
def test_dropin_beamspace_matrix_path(M, L, Nb):
    from cos_amd import getCaponCUDAPy, get_capon
    say("beamspace dropin", M, L, Nb)
    x = uniform_cube(200 + Nb, 3, 12, M)
    V = np.zeros((Nb, L))
    z, R, w, e = getCaponCUDAPy(x, 0.2, L, 1, V, False, False)
    assert e == 0
    ref = O.capon_beamspace_matrix(x.astype(np.complex128), float(np.float32(0.2)), L, 1, Nb)
    assert np.abs(z.ravel() - ref["z"]).max() <= 5e-4 * np.abs(ref["z"]).max()
    assert np.abs(w.reshape(-1, L) - ref["w"]).max() <= 5e-4 * np.abs(ref["w"]).max()
    P = get_capon(x, 0.2, L, 1, Nb=Nb)["P"].ravel()
    assert rel(P, ref["P"]).max() <= 1e-4


# ------------------------------------------------------------------ components

@pytest.mark.parametrize("M,L", [(32, 8), (72, 37)])
def test_buildr_component_elementspace(M, L):
    from cos_amd import BuildR
    say("buildR", M, L)
    x = uniform_cube(5, 3, 10, M)
    Nx, Nz, M = x.shape
    Y, d = 1, 0.2
    R = np.zeros((Nx, Nz - 2 * Y, L, L), np.complex64)
    assert BuildR().getR(R, x, d, L, Y, M, Nx, Nz, 0) == 0
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)["R"]
    assert np.abs(R - ref).max() <= 1e-5 * np.abs(ref).max()


def test_buildr_component_beamspace_input():
    from cos_amd import BuildR
    say("buildR Nb", 16, 8, 3)
    rng = np.random.default_rng(9)
    Nx, Nz, M, L, Nb, Y, d = 2, 7, 16, 8, 3, 1, 0.3
    K = M - L + 1
    xbs = (rng.standard_normal((Nx, Nz, K, Nb)) + 1j * rng.standard_normal((Nx, Nz, K, Nb))).astype(np.complex64)
    R = np.zeros((Nx, Nz - 2 * Y, Nb, Nb), np.complex64)
    assert BuildR().getR(R, xbs, d, L, Y, M, Nx, Nz, Nb) == 0
    for ix in range(Nx):
        for j in range(Nz - 2 * Y):
            S = np.zeros((Nb, Nb), complex)
            for t in range(j, j + 2 * Y + 1):
                for k in range(K):
                    v = xbs[ix, t, k].astype(complex)
                    S += np.outer(v, v.conj())
            S += d * np.trace(S).real / Nb * np.eye(Nb)
            np.testing.assert_allclose(R[ix, j], S.conj(), rtol=1e-5, atol=1e-5 * np.abs(S).max())


@pytest.mark.parametrize("N", [1, 3, 10, 32, 72])
def test_solver_component(N):
    from cos_amd import Solver, solveGaussJordanGPU
    say("solver", N)
    rng = np.random.default_rng(N)
    batch = 5
    A = (rng.standard_normal((batch, N, N)) + 1j * rng.standard_normal((batch, N, N)) + 10 * np.eye(N)).astype(np.complex64)
    b = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    x = np.zeros((batch, N), np.complex64)
    assert Solver().solve(x, A, b, N, batch) == 0
    # column-major A: solve A^T x = b with the C-order arrays (testRTC.pyx:215-236)
    ref = np.linalg.solve(np.swapaxes(A, 1, 2).astype(complex), b.astype(complex)[..., None])[..., 0]
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-4 * np.abs(ref).max())
    x1 = solveGaussJordanGPU(A[0], b[0])
    np.testing.assert_allclose(x1[0], ref[0], rtol=0, atol=1e-4 * np.abs(ref[0]).max())


@pytest.mark.parametrize("N,batch", [(2, 3000), (17, 640), (64, 100), (72, 200)])
def test_solver_matches_oracle_with_exchanges(N, batch):
    """Systems whose diagonal is small against the rest of their column, so that most steps
    exchange rows, against the oracle's restatement of zfsolve_batch (pivot |re| + |im|),
    over a batch of many waves."""
    from cos_amd import Solver
    from oracle import solver_oracle as SO
    say("solver exchanges", N, batch)
    rng = np.random.default_rng(40 + N)
    A = rng.standard_normal((batch, N, N)) + 1j * rng.standard_normal((batch, N, N))
    A[:, np.arange(N), np.arange(N)] *= 1e-3
    A = A.astype(np.complex64)
    b = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    Acm = np.ascontiguousarray(np.swapaxes(A, 1, 2))
    x = np.zeros((batch, N), np.complex64)
    assert Solver().solve(x, Acm, b, N, batch) == 0
    ref, piv = SO.solve_batch(Acm.reshape(batch, -1), b, return_pivots=True)
    assert (piv != np.arange(N)).mean() > 0.5 or N == 2
    # residual relative to |A| |x| + |b|, the backward error of a pivoted elimination
    r = np.abs(np.einsum("sij,sj->si", A.astype(complex), x.astype(complex)) - b)
    scale = np.einsum("sij,sj->si", np.abs(A), np.abs(x)) + np.abs(b)
    assert (r / scale).max() <= 2e-5 * N
    err = np.abs(x - ref).max(axis=1) / np.abs(ref).max(axis=1)
    assert np.median(err) <= 1e-4


@pytest.mark.parametrize("R_on_gpu,x_on_gpu", [(True, True), (True, False), (False, True)])
def test_buildr_device_pointers(R_on_gpu, x_on_gpu):
    """BuildR::getR with device buffers (RTC/BuildR.cpp:59-103): R and/or x as device
    addresses of torch tensors give the host call's R, element space and beamspace input."""
    import torch
    from cos_amd import BuildR
    say("buildR on gpu", R_on_gpu, x_on_gpu)
    x = uniform_cube(6, 3, 12, 24)
    Nx, Nz, M = x.shape
    L, Y, d = 9, 1, 0.2
    shape = (Nx, Nz - 2 * Y, L, L)
    xd = torch.from_numpy(x).cuda()
    Rd = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
    Rh = np.zeros(shape, np.complex64)
    R_arg = Rd.data_ptr() if R_on_gpu else Rh
    x_arg = xd.data_ptr() if x_on_gpu else x
    assert BuildR().getR(R_arg, x_arg, d, L, Y, M, Nx, Nz, 0, R_on_gpu=R_on_gpu, x_on_gpu=x_on_gpu) == 0
    R = Rd.cpu().numpy() if R_on_gpu else Rh
    ref = O.reference_gpu_outputs(x.astype(np.complex128), float(np.float32(d)), L, Y)["R"]
    assert np.abs(R - ref).max() <= 1e-5 * np.abs(ref).max()
    # beamspace-transformed input [Nx][Nz][K][Nb] (buildR_full, loading d tr/Nb)
    rng = np.random.default_rng(10)
    K, Nb = M - L + 1, 4
    xbs = (rng.standard_normal((Nx, Nz, K, Nb)) + 1j * rng.standard_normal((Nx, Nz, K, Nb))).astype(np.complex64)
    Rb = torch.empty((Nx, Nz - 2 * Y, Nb, Nb), dtype=torch.complex64, device="cuda")
    Rbh = np.zeros((Nx, Nz - 2 * Y, Nb, Nb), np.complex64)
    xb_arg = torch.from_numpy(xbs).cuda() if x_on_gpu else xbs
    assert BuildR().getR(Rb.data_ptr() if R_on_gpu else Rbh, xb_arg.data_ptr() if x_on_gpu else xb_arg, d, L, Y,
                         M, Nx, Nz, Nb, R_on_gpu=R_on_gpu, x_on_gpu=x_on_gpu) == 0
    Rb = Rb.cpu().numpy() if R_on_gpu else Rbh
    Rbh2 = np.zeros_like(Rbh)
    assert BuildR().getR(Rbh2, xbs, d, L, Y, M, Nx, Nz, Nb) == 0
    np.testing.assert_array_equal(Rb, Rbh2)
    for ix, j in [(0, 0), (2, Nz - 2 * Y - 1)]:
        v = xbs[ix, j:j + 2 * Y + 1].reshape(-1, Nb).astype(complex)
        S = v.T @ v.conj()
        S += d * np.trace(S).real / Nb * np.eye(Nb)
        np.testing.assert_allclose(Rb[ix, j], np.conj(S), rtol=0, atol=1e-5 * np.abs(S).max())


@pytest.mark.parametrize("flags", [(1, 1, 1), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1)])
def test_solver_device_pointers(flags):
    """Solver::solve with device buffers (RTC/Solver.cpp:59-113) for every combination of
    x_on_gpu, A_on_gpu, b_on_gpu, against LAPACK on the column-major systems."""
    import torch
    from cos_amd import Solver
    xg, Ag, bg = (bool(f) for f in flags)
    say("solver on gpu", flags)
    rng = np.random.default_rng(sum(flags) * 3 + flags[0])
    N, batch = 40, 7
    A = (rng.standard_normal((batch, N, N)) + 1j * rng.standard_normal((batch, N, N)) + 12 * np.eye(N)).astype(np.complex64)
    b = (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))).astype(np.complex64)
    Ad, bd = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
    xd = torch.full((batch, N), float("nan"), dtype=torch.complex64, device="cuda")
    xh = np.zeros((batch, N), np.complex64)
    assert Solver().solve(xd.data_ptr() if xg else xh, Ad.data_ptr() if Ag else A, bd.data_ptr() if bg else b, N,
                          batch, x_on_gpu=xg, A_on_gpu=Ag, b_on_gpu=bg) == 0
    x = xd.cpu().numpy() if xg else xh
    ref = np.linalg.solve(np.swapaxes(A, 1, 2).astype(complex), b.astype(complex)[..., None])[..., 0]
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-4 * np.abs(ref).max())


def test_out_of_memory_is_reported_and_the_plan_survives():
    """A request past the device's memory (RTC/Capon.cpp:116-133 reports
    cudaErrorMemoryValueTooLarge) returns the out-of-memory code with a message, and the
    same process keeps working afterwards."""
    import torch
    from cos_amd import getCaponCUDAPy
    from cos_amd._lib import lib
    say("out of memory")
    free, total = torch.cuda.mem_get_info()
    # R of 6000 x 4092 pixels at L = 64 is 804 GB, past the device; the cube itself (12.6 GB)
    # is allocated first, then R fails before anything is copied from the host pointers
    M, L, Y, Nx, Nz = 64, 64, 2, 6000, 4096
    need_R = Nx * (Nz - 2 * Y) * L * L * 8
    assert need_R > 2 * total
    h = np.zeros(1, np.complex64)
    code = lib().cos_capon_get_capon_p(None, None, h.ctypes.data, None, h.ctypes.data, 0.01, L, Y, M, Nx, Nz, 0)
    assert code == 2, code                      # hipErrorOutOfMemory
    assert b"device buffer" in lib().cos_capon_last_error()
    # the drop-in still works afterwards
    xs = uniform_cube(7, 2, 12, 16)
    z2, R2, w2, e = getCaponCUDAPy(xs, 0.1, 8, 1, np.array([0]), False, False)
    assert e == 0
    ref = O.reference_gpu_outputs(xs.astype(np.complex128), float(np.float32(0.1)), 8, 1)
    assert np.abs(z2 - ref["z"]).max() <= 2e-4 * np.abs(ref["z"]).max() + 1e-6


# ------------------------------------------------------------------ edge cases

def test_single_output_row_and_single_beam():
    from cos_amd import get_capon
    say("edge single row")
    x = uniform_cube(11, 1, 5, 16)
    o = get_capon(x, 0.1, 8, 2)
    ref = O.capon(x.astype(np.complex128), float(np.float32(0.1)), 8, 2)
    assert o["P"].shape == (1, 1)
    assert rel(o["P"].ravel(), ref["P"]).max() < 1e-5


def test_zero_cube_is_finite_with_loading():
    from cos_amd import get_capon
    say("edge zero cube")
    x = np.zeros((2, 6, 16), np.complex64)
    x[:, :, 3] = 1.0        # rank-deficient: loading keeps it PD
    o = get_capon(x, 0.1, 8, 1)
    assert np.all(np.isfinite(o["P"])) and np.all(o["P"] > 0)
    ref = O.capon(x.astype(np.complex128), float(np.float32(0.1)), 8, 1)
    assert rel(o["P"].ravel(), ref["P"]).max() < 1e-4


def test_full_frame_properties():
    """Headline frame (256 x 2048 x 64, L = 32): size-independent checks on every pixel
    plus oracle parity on a strided sample."""
    import torch
    from cos_amd import synth
    say("full frame cfg2")
    x = synth.config2()
    o, _ = _plan_run(x, 0.01, 32, 2)
    P, w, z = o["P"], o["w"], o["z"]
    assert np.all(np.isfinite(P)) and np.all(P > 0)
    s = w.sum(axis=1)                              # distortionless a^H w = 1
    assert np.abs(s - 1).max() < 1e-3
    ix = np.arange(0, 256, 37)
    sub = c_capon(x, 0.01, 32, 2, ix=(0, 256), j=(1000, 1010), want_w=False)
    Pn = P.reshape(256, 2044)[:, 1000:1010]
    assert rel(Pn, sub["P"]).max() <= 1e-4


def test_graph_replay_matches_eager():
    """The frame captured in a hipGraph and replayed (with eager work and synchronisation
    between replays) gives the eager result every time; new pointers re-capture."""
    import torch
    from cos_amd import CaponPlan, synth
    say("graph replay")
    x = torch.from_numpy(synth.config2(Nx=8, Nz=256, M=64, seed=2)).cuda()
    plan = CaponPlan()
    eager = plan.run(x, 0.01, 32, 2, precision="fp32+fallback")
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in eager.items()}
    out = {k: torch.full_like(v, float("nan")) if v.is_floating_point() or v.is_complex() else torch.zeros_like(v)
           for k, v in eager.items()}
    for rep in range(4):
        for v in out.values():
            v.zero_()
        plan.run(x, 0.01, 32, 2, precision="fp32+fallback", out=out, graph=True)
        torch.cuda.synchronize()
        _ = (x.abs().sum() * 2).item()                      # eager work between replays
        for k in ref:
            assert torch.equal(out[k], ref[k]), (rep, k)
    # different output tensors -> a new capture, same values
    out2 = {k: torch.empty_like(v) for k, v in eager.items()}
    plan.run(x, 0.01, 32, 2, precision="fp32+fallback", out=out2, graph=True)
    torch.cuda.synchronize()
    for k in ref:
        assert torch.equal(out2[k], ref[k]), k
    # rows with tiny loading: the fallback pass runs inside the graph
    rows = {"L": torch.full((256,), 32, dtype=torch.int32).cuda(), "M": torch.full((256,), 64, dtype=torch.int32).cuda(),
            "off": torch.zeros(256, dtype=torch.int32).cuda(), "d": torch.full((256,), 1e-6).cuda()}
    e2 = plan.run(x, 0.0, 0, 2, rows=rows, precision="fp32+fallback")
    torch.cuda.synchronize()
    n_eager = plan.fallback_count()
    e2r = {k: v.clone() for k, v in e2.items()}
    o3 = {k: torch.empty_like(v) for k, v in e2.items()}
    for rep in range(3):
        plan.run(x, 0.0, 0, 2, rows=rows, precision="fp32+fallback", out=o3, graph=True)
        torch.cuda.synchronize()
        assert plan.fallback_count() == n_eager == 8 * 252
        for k in e2r:
            assert torch.equal(o3[k], e2r[k]), (rep, k)


def test_graph_replay_follows_row_tables_changed_in_place():
    """The lane-shape split of a table-driven frame is planned on the device inside the
    frame graph, so a replay after the caller rewrites the L, M and d tables in place
    (same pointers, same Lmax) computes the new frame: rows moved between lane shapes
    (L 12 <-> 40 <-> 60) and into the fallback match an eager run on the new tables."""
    import torch
    from cos_amd import CaponPlan
    say("graph replay, tables changed in place")
    Nx, Nz, Mmax, Y = 4, 300, 128, 2
    x = torch.from_numpy(uniform_cube(44, Nx, Nz, Mmax)).cuda()
    rng = np.random.default_rng(45)

    def tables(seed):
        r = np.random.default_rng(seed)
        L = r.choice([12, 24, 40, 60], Nz).astype(np.int32)
        M = np.minimum(2 * L + r.integers(0, 8, Nz), Mmax).astype(np.int32)
        off = ((Mmax - M) // 2).astype(np.int32)
        d = np.where(r.random(Nz) < 0.1, 1e-6, 0.01).astype(np.float32)
        return L, M, off, d

    rows = {k: torch.from_numpy(v).cuda() for k, v in zip(("L", "M", "off", "d"), tables(1))}
    plan = CaponPlan()
    out = {k: torch.empty(s, dtype=t, device="cuda") for k, s, t in
           (("P", (Nx * (Nz - 2 * Y),), torch.float32), ("z", (Nx * (Nz - 2 * Y),), torch.complex64),
            ("w", (Nx * (Nz - 2 * Y), 60), torch.complex64))}
    for seed in (1, 2, 3):
        for k, v in zip(("L", "M", "off", "d"), tables(seed)):
            rows[k].copy_(torch.from_numpy(v))                     # in place: same pointers
        for v in out.values():
            v.zero_()
        plan.run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", out=out, graph=True)
        torch.cuda.synchronize()
        n_graph = plan.fallback_count()
        eager = CaponPlan().run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback")
        torch.cuda.synchronize()
        for k in out:
            assert torch.equal(out[k][:, :eager[k].shape[-1]] if k == "w" else out[k], eager[k]), (seed, k)
        assert bool(torch.isfinite(out["P"]).all())
        Lh = rows["L"].cpu().numpy()
        c = np.arange(Nz - 2 * Y) + Y
        assert n_graph == Nx * int((rows["d"].cpu().numpy()[c] < 1e-3).sum()), seed
        sel = np.arange(0, Nx * (Nz - 2 * Y), 5)
        ref = O.capon_rows(x.cpu().numpy().astype(np.complex128), rows["d"].cpu().numpy().astype(np.float64), Lh,
                           rows["M"].cpu().numpy(), rows["off"].cpu().numpy(), Y,
                           pixels=(sel // (Nz - 2 * Y), sel % (Nz - 2 * Y)))
        assert rel(out["P"].cpu().numpy()[sel], ref["P"]).max() <= 1e-4, seed
    del rng


def test_graph_replay_after_plan_buffers_grow():
    """Graph of frame A (row tables, fallback on), then an eager larger frame B on the same
    plan, whose fallback buffers (flags, list, count) must grow, then graph A again: the
    second graph call re-captures onto the new buffers and matches eager A exactly, with
    the fallback count of A."""
    import torch
    from cos_amd import CaponPlan, synth
    say("graph replay after regrow")
    def rows_for(Nz, d):
        return {"L": torch.full((Nz,), 32, dtype=torch.int32).cuda(), "M": torch.full((Nz,), 64, dtype=torch.int32).cuda(),
                "off": torch.zeros(Nz, dtype=torch.int32).cuda(), "d": torch.full((Nz,), d).cuda()}
    xa = torch.from_numpy(synth.config2(Nx=4, Nz=128, M=64, seed=2)).cuda()
    xb = torch.from_numpy(synth.config2(Nx=16, Nz=512, M=64, seed=3)).cuda()
    ra, rb = rows_for(128, 1e-6), rows_for(512, 1e-6)
    plan = CaponPlan()
    ea = plan.run(xa, 0.0, 0, 2, rows=ra, precision="fp32+fallback")
    torch.cuda.synchronize()
    na = plan.fallback_count()
    ea = {k: v.clone() for k, v in ea.items()}
    oa = {k: torch.empty_like(v) for k, v in ea.items()}
    plan.run(xa, 0.0, 0, 2, rows=ra, precision="fp32+fallback", out=oa, graph=True)
    torch.cuda.synchronize()
    for k in ea:
        assert torch.equal(oa[k], ea[k]), ("first graph", k)
    eb = plan.run(xb, 0.0, 0, 2, rows=rb, precision="fp32+fallback")     # grows the plan's buffers
    torch.cuda.synchronize()
    assert plan.fallback_count() == 16 * 508
    for v in oa.values():
        v.zero_()
    plan.run(xa, 0.0, 0, 2, rows=ra, precision="fp32+fallback", out=oa, graph=True)
    torch.cuda.synchronize()
    assert plan.fallback_count() == na == 4 * 124
    for k in ea:
        assert torch.equal(oa[k], ea[k]), ("graph after regrow", k)
    del eb


# ------------------------------------------------------------------ sliding-DFT beamspace

@pytest.mark.parametrize("M,L,Y,Nb,d,Nz", [(12, 6, 1, 2, 0.2, 40), (32, 16, 2, 5, 0.05, 40), (20, 9, 0, 1, 0.1, 40),
                                           (40, 20, 3, 8, 0.01, 40), (16, 8, 1, 4, 0.1, 40),
                                           # K = 1 (no sliding step); several blocks, the last one
                                           # with a wave wholly past the frame, odd L and K
                                           (8, 8, 1, 3, 0.1, 40), (50, 23, 2, 5, 0.02, 300)])
def test_beamspace_sliding_matches_oracle(M, L, Y, Nb, d, Nz):
    import torch
    from cos_amd import CaponPlan
    say("beamspace sliding", M, L, Y, Nb, d, Nz)
    x = uniform_cube(300 + Nb + M, 3, Nz, M)
    o = CaponPlan().run_beamspace(torch.from_numpy(x).cuda(), d, L, Y, Nb, want=("P", "z", "w_bs", "w"))
    torch.cuda.synchronize()
    ref = O.capon_beamspace(x.astype(np.complex128), d, L, Y, Nb)
    P = o["P"].cpu().numpy()
    assert rel(P, ref["P"]).max() <= 1e-4
    z = o["z"].cpu().numpy()
    assert np.abs(z - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max()
    wbs = o["w_bs"].cpu().numpy()
    assert np.abs(wbs - ref["w_bs"]).max() <= 1e-4 * np.abs(ref["w_bs"]).max()
    w = o["w"].cpu().numpy()
    assert np.abs(w - ref["w"]).max() <= 1e-4 * np.abs(ref["w"]).max()


def test_beamspace_sonar_slice():
    """The 96-element, L = 48, Nb = 5, Yavg = 4 sonar geometry on 8 look directions, power
    only (the kernel variant that accumulates no subarray means)."""
    import torch
    from cos_amd import CaponPlan, synth
    say("beamspace sonar slice", 8, 1024, 96, 48, 5)
    x = synth.config4(Nx=8, Nz=1024, M=96, seed=4)
    o = CaponPlan().run_beamspace(torch.from_numpy(x).cuda(), 0.01, 48, 4, 5, want=("P",))
    torch.cuda.synchronize()
    P = o["P"].cpu().numpy()
    ix, j = O.all_pixels(8, 1024, 4)
    sel = np.arange(0, len(ix), 7)
    ref = O.capon_beamspace(x.astype(np.complex128), 0.01, 48, 4, 5, pixels=(ix[sel], j[sel]))
    e = rel(P[sel], ref["P"])
    print(f"sonar P rel err max {e.max():.3e}", flush=True)
    assert e.max() <= 1e-4


def test_beamspace_sonar_full_frame():
    """The whole sonar frame (512 look directions x 8192 ranges x 96 elements, L = 48,
    Nb = 5, Yavg = 4; 4,190,208 pixels): every power finite and positive, the frame
    deterministic across two runs, and a strided sample of 512 pixels (every 8183rd, all
    beams and depths) against the FP64 oracle within 1e-4."""
    import torch
    from cos_amd import CaponPlan, synth
    Nx, Nz, M, L, Y, Nb, d = 512, 8192, 96, 48, 4, 5, 0.01
    say("beamspace sonar full frame", Nx, Nz, M, L, Nb, Y)
    x = synth.config4_torch(Nx, Nz, M, seed=4, device="cuda")
    plan = CaponPlan()
    o1 = plan.run_beamspace(x, d, L, Y, Nb, want=("P",))["P"].clone()
    o2 = plan.run_beamspace(x, d, L, Y, Nb, want=("P",))["P"]
    torch.cuda.synchronize()
    npx = Nx * (Nz - 2 * Y)
    assert o1.numel() == npx == 4190208
    assert bool(torch.isfinite(o1).all()) and bool((o1 > 0).all())
    assert torch.equal(o1, o2)
    sel = np.arange(0, npx, 8183)
    Nzp = Nz - 2 * Y
    ib, jb = sel // Nzp, sel % Nzp
    P = o1.cpu().numpy()[sel]
    ref = np.empty(len(sel))
    for b in np.unique(ib):                      # the oracle on each sampled beam's rows
        k = np.nonzero(ib == b)[0]
        xb = x[int(b):int(b) + 1].cpu().numpy().astype(np.complex128)
        ref[k] = O.capon_beamspace(xb, d, L, Y, Nb, pixels=(np.zeros(len(k), int), jb[k]))["P"]
    e = rel(P, ref)
    print(f"sonar full frame: P rel err max {e.max():.3e} over {len(sel)} sampled pixels", flush=True)
    assert e.max() <= 1e-4


def test_expanding_aperture_full_frame():
    """The whole expanding-aperture frame (256 beams x 4096 ranges, M 16..128, L = M/2,
    5% of rows at d = 1e-5; 1,047,552 pixels) in FP32 with the FP64 fallback, against the
    all-FP64 run of the same frame on every pixel (P within 1e-4), every weight vector
    distortionless (sum w = 1, a = 1), the fallback count that of the flagged rows, and a
    strided sample of 256 pixels against the FP64 oracle."""
    import torch
    from cos_amd import CaponPlan, synth
    Nx, Nz, Mmax, Y = 256, 4096, 128, 2
    say("expanding aperture full frame", Nx, Nz, Mmax)
    x, Mz, Lz, off, dz = synth.config3_torch(Nx, Nz, seed=3, Mmax=Mmax, device="cuda")
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    plan = CaponPlan()
    o = plan.run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", want=("P", "w", "flags"))
    o64 = CaponPlan().run(x, 0.0, 0, Y, rows=rows, precision="fp64", want=("P",))
    torch.cuda.synchronize()
    Nzp = Nz - 2 * Y
    npx = Nx * Nzp
    c = np.arange(Nzp) + Y
    assert o["P"].numel() == npx == 1047552
    e = (o["P"].double() - o64["P"].double()).abs() / o64["P"].double()
    print(f"aperture full frame vs all-FP64: P max {e.max().item():.3e} mean {e.mean().item():.3e}", flush=True)
    assert e.max().item() <= 1e-4
    # flagged: the condition bound 1 + L/d above 1e6 in FP32 (the rows at d = 1e-5 with
    # L > 10; shallower ones stay in FP32)
    ill = (np.float32(1) + Lz[c].astype(np.float32) / dz[c]) > np.float32(1e6)
    assert 0 < ill.sum() < (dz[c] < 1e-3).sum()
    assert plan.fallback_count() == Nx * int(ill.sum())
    fl = o["flags"].reshape(Nx, Nzp).cpu().numpy()
    assert ((fl & 1) != 0)[:, ill].all() and (fl[:, ~ill] == 0).all()
    # distortionless: the first L(z) weights of every pixel sum to 1, the rest are 0
    w = o["w"].reshape(Nx, Nzp, 64)
    Lc = torch.from_numpy(Lz[c].astype(np.int64)).cuda()
    mask = (torch.arange(64, device="cuda")[None, :] < Lc[:, None])[None]
    s = (w * mask).sum(-1)
    assert (s - 1).abs().max().item() <= 1e-3
    sel = np.arange(0, npx, 4099)
    ref = O.capon_rows(x.cpu().numpy().astype(np.complex128), dz.astype(np.float64), Lz, Mz, off, Y,
                       pixels=(sel // Nzp, sel % Nzp))
    eo = rel(o["P"].cpu().numpy()[sel], ref["P"])
    print(f"aperture full frame vs oracle: P max {eo.max():.3e} over {len(sel)} pixels", flush=True)
    assert eo.max() <= 1e-4


# ------------------------------------------------------------------ reference options

OPTION_CASES = [
    # (M, L, Y, precision): 16-lane one column, two columns, 64-lane, FP64 direct, FP64 wide
    (24, 12, 1, "fp32"), (48, 24, 2, "fp32"), (80, 40, 1, "fp32"), (48, 24, 2, "fp64"), (72, 40, 1, "fp64"),
]


@pytest.mark.parametrize("M,L,Y,prec", OPTION_CASES)
@pytest.mark.parametrize("fb,zm,pm", [(True, "amplitude", "loaded"), (False, "power", "unloaded"),
                                      (True, "power", "unloaded")])
def test_reference_options_match_oracle(M, L, Y, prec, fb, zm, pm):
    """Forward-backward averaging, power-Capon amplitude and the unloaded power against the
    FP64 oracle: P and z within 1e-4 (FP32) or 2e-6 (FP64 stored as float32), w within 1e-3
    of its largest entry (FP32)."""
    say("options", M, L, Y, prec, fb, zm, pm)
    x = uniform_cube(M + 7 * L, 3, 12, M)
    o, _ = _plan_run(x, 0.05, L, Y, precision=prec, fb_avg=fb, z_mode=zm, p_mode=pm)
    ref = O.capon_options(x.astype(np.complex128), 0.05, L, Y, fb=fb,
                          z_mode=1 if zm == "power" else 0, p_mode=1 if pm == "unloaded" else 0)
    tol = 1e-4 if prec == "fp32" else 2e-6
    eP = rel(o["P"], ref["P"])
    ez = np.abs(o["z"] - ref["z"]) / np.abs(ref["z"]).max()
    print(f"P err {eP.max():.2e} z err {ez.max():.2e}", flush=True)
    assert eP.max() <= tol
    assert ez.max() <= tol
    np.testing.assert_allclose(o["w"].reshape(-1, L), ref["w"], atol=(1e-3 if prec == "fp32" else 1e-5) * np.abs(ref["w"]).max())


@pytest.mark.parametrize("M,L,Y,Nb,fb", [(24, 12, 1, 3, False), (48, 24, 2, 6, True), (64, 40, 1, 5, False)])
def test_subspace_matches_oracle(M, L, Y, Nb, fb):
    """Capon in the subspace of a random V against the FP64 oracle (FP32 covariance, FP64
    subspace solve): P within 1e-4, z and w within 1e-4 of their scale."""
    import torch
    from cos_amd import CaponPlan
    say("subspace", M, L, Y, Nb, fb)
    x = uniform_cube(M + L + Nb, 3, 14, M)
    rng = np.random.default_rng(Nb)
    V = ((rng.standard_normal((Nb, L)) + 1j * rng.standard_normal((Nb, L))) / np.sqrt(L)).astype(np.complex64)
    o = CaponPlan().run_subspace(torch.from_numpy(x).cuda(), 0.05, L, Y, torch.from_numpy(V).cuda(), fb_avg=fb)
    torch.cuda.synchronize()
    o = {k: v.cpu().numpy() for k, v in o.items()}
    ref = O.capon_subspace(x.astype(np.complex128), 0.05, L, Y, V.astype(np.complex128), fb=fb)
    eP = rel(o["P"], ref["P"])
    print(f"subspace P err {eP.max():.2e}", flush=True)
    assert eP.max() <= 1e-4
    assert np.abs(o["z"] - ref["z"]).max() <= 1e-4 * np.abs(ref["z"]).max()
    assert np.abs(o["w"] - ref["w"]).max() <= 1e-4 * np.abs(ref["w"]).max()


def test_subspace_butler_equals_beamspace():
    """The Butler rows as V reproduce the sliding-DFT beamspace path (loading in beamspace)."""
    import torch
    from cos_amd import CaponPlan
    say("subspace butler", 32, 16, 1, 5)
    x = uniform_cube(77, 2, 20, 32)
    B = O.butler_matrix(5, 16).astype(np.complex64)
    plan = CaponPlan()
    xt = torch.from_numpy(x).cuda()
    s = plan.run_subspace(xt, 0.05, 16, 1, torch.from_numpy(B).cuda())
    b = plan.run_beamspace(xt, 0.05, 16, 1, 5, want=("P", "z", "w_bs"))
    torch.cuda.synchronize()
    assert rel(s["P"].cpu().numpy(), b["P"].cpu().numpy()).max() <= 2e-4
    np.testing.assert_allclose(s["w"].cpu().numpy(), b["w_bs"].cpu().numpy(), atol=2e-4 * float(b["w_bs"].abs().max()))
