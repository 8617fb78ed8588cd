"""Pins the oracle (CPU, no GPU).

Each oracle function is checked against the reference's own known-answer vectors
(tests/golden/reference_fixtures.json, parsed from apps/RealTimeCaponTest/TestData.h
and framework/linalg/tests/testMatrixOperations.py), against brute force on tiny
inputs in the manner of the reference's testBuildRGPU
(framework/beamformer/capon/tests/testRTC.pyx:133-163), against LAPACK through numpy,
or against closed forms -- never against a retyped copy of its own formula.
"""
import numpy as np
import pytest

from oracle import c_capon
from oracle import capon_oracle as O


def cvec(v):
    a = np.asarray(v, dtype=np.float64)
    return a[..., 0] + 1j * a[..., 1]


# ---------------------------------------------------------------- TestData.h vectors

def test_testdata_R8_unnormalised(golden):
    x = cvec(golden["testX16"])                      # 1..16
    rows = x[None, None, :]                          # one pixel, Y = 0
    S = O.covariance_unnormalised(rows, 8)[0]
    np.testing.assert_array_equal(S.real, np.array(golden["R8_unnormalised"]))
    np.testing.assert_array_equal(S.imag, 0)


def test_testdata_R8_normalised(golden):
    x = cvec(golden["testX16"])
    S = O.covariance_unnormalised(x[None, None, :], 8)[0]
    K = 16 - 8 + 1
    np.testing.assert_allclose((S / K).real, np.array(golden["R8_normalised_4dp"]), atol=6e-5)


def test_testdata_R8_yavg1_loaded(golden):
    # X = [testRightSide testX16 testX16], one pixel with Yavg = 1, L = 8.  The loading
    # factor is not printed in TestData.h; d = 0.8 is the value its diagonal implies.
    ones = cvec(golden["testRightside"])
    x16 = cvec(golden["testX16"])
    cube = np.stack([ones, x16, x16])[None]          # [1][3][16]
    rows = O.window_rows(cube, np.array([0]), np.array([0]), 1)
    S = O.covariance_unnormalised(rows, 8)
    SL = O.diagonal_load(S, 0.8, 8)[0]
    np.testing.assert_allclose(SL.real, np.array(golden["R8_yavg1_loaded_4dp"]), atol=0.051)
    # off-diagonals do not depend on the loading: they pin the temporal sum alone
    off = ~np.eye(8, dtype=bool)
    np.testing.assert_allclose(S[0].real[off], np.array(golden["R8_yavg1_loaded_4dp"])[off], atol=0.051)


def test_testdata_R16_unnormalised(golden):
    x = cvec(golden["testX32_active"])[:16]
    x32 = np.concatenate([x, x])                     # the [1..16, 1..16] input of TestData.h:123-173
    S = O.covariance_unnormalised(x32[None, None, :], 16)[0]
    np.testing.assert_array_equal(S.real, np.array(golden["R16_unnormalised"]))


def test_gpu_R_layout_is_conjugate_column_major():
    # memory[r*L + c] = conj(S[r][c]): reading it column-major gives S (testRTC.pyx:180).
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 3, 10)) + 1j * rng.standard_normal((1, 3, 10))
    o = O.capon(x, 0.2, 4, 1, want_R=True)
    Rmem = o["R"][0]
    rows = O.window_rows(x, np.array([0]), np.array([0]), 1)
    SL = O.diagonal_load(O.covariance_unnormalised(rows, 4), 0.2, 4)[0]
    np.testing.assert_allclose(Rmem.reshape(-1).reshape(4, 4, order="F"), SL, atol=1e-12)
    np.testing.assert_allclose(Rmem, np.conj(SL), atol=1e-12)


# ---------------------------------------------------------------- brute-force covariance

def test_covariance_brute_force():
    """Explicit loops over subarrays and snapshots, as testBuildRGPU builds its answer."""
    rng = np.random.default_rng(1)
    Nx, Nz, M, L, Y, d = 2, 7, 10, 4, 1, 0.2
    x = rng.standard_normal((Nx, Nz, M)) + 1j * rng.standard_normal((Nx, Nz, M))
    K = M - L + 1
    for ix in range(Nx):
        for j in range(Nz - 2 * Y):
            R = np.zeros((L, L), complex)
            for t in range(j, j + 2 * Y + 1):
                for k in range(K):
                    u = x[ix, t, k:k + L]
                    R += np.outer(u, u.conj())
            R += d * np.trace(R).real / L * np.eye(L)
            rows = O.window_rows(x, np.array([ix]), np.array([j]), Y)
            got = O.diagonal_load(O.covariance_unnormalised(rows, L), d, L)[0]
            np.testing.assert_allclose(got, R, rtol=1e-13, atol=1e-12)


# ---------------------------------------------------------------- Cholesky / solves

def test_cholesky_real_fixture(golden):
    lg = golden["linalg"]
    A, R = cvec(lg["A"]), cvec(lg["R"])
    np.testing.assert_allclose(O.cholesky_upper(A), R, atol=1e-14)
    np.testing.assert_allclose(O.solve_cholesky(A, cvec(lg["b1"])), cvec(lg["x1"]), atol=1e-14)
    np.testing.assert_allclose(O.solve_cholesky(A, cvec(lg["b2"])), cvec(lg["x2"]), atol=1e-14)


def test_cholesky_complex_fixture(golden):
    lg = golden["linalg"]
    A, R = cvec(lg["complexA"]), cvec(lg["complexR"])
    b, y, x = cvec(lg["complexb"]), cvec(lg["complexy"]), cvec(lg["complexx"])
    U = O.cholesky_upper(A)
    np.testing.assert_allclose(U, R, atol=1e-14)
    np.testing.assert_allclose(O.forward_solve(U.conj().T, b), y, atol=1e-14)
    np.testing.assert_allclose(O.back_solve(U, y), x, atol=1e-14)
    np.testing.assert_allclose(O.solve_cholesky(A, b), x, atol=1e-14)


def test_cholesky_reads_upper_triangle_only():
    rng = np.random.default_rng(2)
    B = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    A = B @ B.conj().T + 6 * np.eye(6)
    junk = np.triu(A) + np.tril(rng.standard_normal((6, 6)), -1) * 1e3
    np.testing.assert_allclose(O.cholesky_upper(junk), np.linalg.cholesky(A).conj().T, atol=1e-12)


def test_solve_matches_lapack_batched():
    rng = np.random.default_rng(3)
    B = rng.standard_normal((5, 8, 8)) + 1j * rng.standard_normal((5, 8, 8))
    A = B @ np.conj(np.swapaxes(B, -1, -2)) + np.eye(8)
    b = rng.standard_normal((5, 8)) + 1j * rng.standard_normal((5, 8))
    np.testing.assert_allclose(O.solve_cholesky(A, b), np.linalg.solve(A, b[..., None])[..., 0], rtol=1e-11)


# ---------------------------------------------------------------- Capon outputs

def test_capon_identities_and_lapack():
    rng = np.random.default_rng(4)
    Nx, Nz, M, L, Y, d = 2, 9, 12, 5, 2, 0.1
    x = rng.standard_normal((Nx, Nz, M)) + 1j * rng.standard_normal((Nx, Nz, M))
    o = O.capon(x, d, L, Y)
    ix, j = O.all_pixels(Nx, Nz, Y)
    rows = O.window_rows(x, ix, j, Y)
    K = M - L + 1
    RL = O.diagonal_load(O.covariance_unnormalised(rows, L) / (K * (2 * Y + 1)), d, L)
    a = np.ones(L)
    Rinv = np.linalg.inv(RL)
    P = 1.0 / np.einsum("l,blm,m->b", a, Rinv, a).real
    np.testing.assert_allclose(o["P"], P, rtol=1e-11)
    np.testing.assert_allclose(o["w"].sum(-1), 1.0, atol=1e-12)                       # a^H w = 1
    np.testing.assert_allclose(np.einsum("bi,bij,bj->b", o["w"].conj(), RL, o["w"]).real, P, rtol=1e-10)
    # amplitude: sliding subarray sums of the centre snapshot / K
    for n in range(3):
        c = rows[n, Y]
        g = np.array([c[l:l + K].sum() for l in range(L)]) / K
        np.testing.assert_allclose(o["z"][n], np.vdot(o["w"][n], g), rtol=1e-12)


def test_capon_white_noise_closed_form():
    """Three orthogonal snapshots (K = 1) give S = 2 (I - q q^T) for the missing
    direction q; R_L^-1 then has a closed form (Sherman-Morrison)."""
    L, M = 4, 4                                   # K = 1 subarray
    Y = 1
    Q = np.linalg.qr(np.random.default_rng(5).standard_normal((4, 4)))[0]
    x = np.zeros((1, 3, M), complex)
    x[0, 0] = Q[0] * np.sqrt(2)
    x[0, 1] = Q[1] * np.sqrt(2)
    x[0, 2] = Q[2] * np.sqrt(2)
    S = O.covariance_unnormalised(O.window_rows(x, np.array([0]), np.array([0]), Y), L)[0]
    np.testing.assert_allclose(S, 2 * (np.eye(4) - np.outer(Q[3], Q[3])), atol=1e-12)
    # closed form with one null direction: R_L = (2/3)(I - q q^T) + d (2*3/4)/3 I
    d = 0.5
    o = O.capon(x, d, L, Y)
    c = 2.0 / 3.0
    mu = d * c * 3 / 4
    q = Q[3]
    a = np.ones(4)
    # R_L = (c+mu) I - c q q^T  ->  R_L^-1 = I/(c+mu) + c q q^T /((c+mu) mu)
    aRa = a @ a / (c + mu) + c * (q @ a) ** 2 / ((c + mu) * mu)
    np.testing.assert_allclose(o["P"][0], 1 / aRa, rtol=1e-12)


def test_zpow_reference_quirk_differs_from_power():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1, 5, 8)) + 1j * rng.standard_normal((1, 5, 8))
    o = O.capon(x, 0.2, 4, 1)
    for n in range(len(o["P"])):
        w = o["w"][n]
        rows = O.window_rows(x, np.array([0]), np.array([n]), 1)
        K = 5
        RL = O.diagonal_load(O.covariance_unnormalised(rows, 4) / (K * 3), 0.2, 4)[0]
        Rup = np.triu(RL)
        quirk = sum(sum(np.conj(w[jj]) * Rup[i, jj] for jj in range(4)) * w[i] for i in range(4))
        assert abs(o["zPow_ref"][n] - quirk) < 1e-12
        assert abs(quirk - o["P"][n]) > 1e-6      # the quirk is not w^H R_L w


def test_c_oracle_matches_numpy_oracle():
    rng = np.random.default_rng(7)
    Nx, Nz, M, L, Y, d = 3, 11, 16, 7, 2, 0.05
    x = (rng.standard_normal((Nx, Nz, M)) + 1j * rng.standard_normal((Nx, Nz, M))).astype(np.complex64)
    o = O.capon(x, d, L, Y)
    c = c_capon(x, d, L, Y)
    np.testing.assert_allclose(c["P"].ravel(), o["P"], rtol=1e-11)
    np.testing.assert_allclose(c["z"].ravel(), o["z"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(c["w"].reshape(-1, L), o["w"], rtol=1e-10, atol=1e-12)
    sub = c_capon(x, d, L, Y, ix=(1, 3), j=(2, 5))
    full = o["P"].reshape(Nx, Nz - 2 * Y)
    np.testing.assert_allclose(sub["P"], full[1:3, 2:5], rtol=1e-11)


# ---------------------------------------------------------------- beamspace

def test_butler_matches_reference_fixtures(golden):
    lg = golden["linalg"]
    np.testing.assert_allclose(O.butler_matrix(3, 3), cvec(lg["but3"]), atol=1e-14)
    # LA/beamspace.py:18-34 is the plain DFT (beams 0..m-1); check the sign/normalisation
    # convention on the 4x4 fixture with beams 0..3 taken directly.
    l = np.arange(4)
    B4 = np.exp(-2j * np.pi * np.outer(l, l) / 4) / 2
    np.testing.assert_allclose(B4, cvec(lg["but4"]), atol=1e-14)
    np.testing.assert_allclose(O.butler_matrix(3, 4)[:2], B4[:2], atol=1e-14)


def test_beam_index_symmetric_map():
    np.testing.assert_array_equal(O.beam_index(5, 48), [0, 1, 2, 46, 47])
    # even Nb: 3 + 16 - 2*2 - 1 = 14, i.e. beam -2; the '+1' that would give -1 never fires
    np.testing.assert_array_equal(O.beam_index(4, 16), [0, 1, 2, 14])


def test_sliding_dft_equals_direct():
    rng = np.random.default_rng(8)
    rows = rng.standard_normal((3, 2, 20)) + 1j * rng.standard_normal((3, 2, 20))
    L, Nb = 8, 5
    v = O.sliding_dft(rows, L, Nb)
    u = O.subarrays(rows, L)
    direct = u @ O.butler_matrix(Nb, L).T
    np.testing.assert_allclose(v, direct, atol=1e-12)


def test_beamspace_full_basis_equals_element_space():
    """With Nb = L distinct beams (odd L) B is unitary: beamspace Capon with loading
    d tr/Nb reproduces element-space P, z and w."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal((1, 7, 12)) + 1j * rng.standard_normal((1, 7, 12))
    L, Y, d = 5, 1, 0.1
    e = O.capon(x, d, L, Y)
    b = O.capon_beamspace(x, d, L, Y, Nb=L)
    np.testing.assert_allclose(b["P"], e["P"], rtol=1e-11)
    np.testing.assert_allclose(b["z"], e["z"], rtol=1e-10)
    np.testing.assert_allclose(b["w"], e["w"], atol=1e-11)
    m = O.capon_beamspace_matrix(x, d, L, Y, Nb=L)
    np.testing.assert_allclose(m["P"], e["P"], rtol=1e-11)
    np.testing.assert_allclose(m["w"], e["w"], atol=1e-11)


def _beamspace_case(seed):
    # Nb < L: B has orthonormal rows but is not unitary, so B S B^H loses information and
    # the place where the loading is applied matters (K (2Y+1) = 27 snapshots >= Nb)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((2, 9, 24)) + 1j * rng.standard_normal((2, 9, 24))
    return x, 16, 1, 5


def test_beamspace_matrix_unloaded_equals_sliding_beamspace():
    """At d = 0 the reference's matrix path (B S B^H by products, a_bs = B 1, y = B^H y_bs,
    then the element-space output stage, RTC/Capon.cpp:285-402) and the sliding-DFT path
    (Gram of the beamspace subarrays, w = B^H w_bs, z = w_bs^H B g, RTC/Capon.cpp:202-227,
    RTC/capon_kernel.cu:162-233) are two computations of one estimator: P, w and z agree."""
    x, L, Y, Nb = _beamspace_case(21)
    m = O.capon_beamspace_matrix(x, 0.0, L, Y, Nb)
    b = O.capon_beamspace(x, 0.0, L, Y, Nb)
    np.testing.assert_allclose(m["P"], b["P"], rtol=1e-10)
    np.testing.assert_allclose(m["w"], b["w"], rtol=0, atol=1e-10 * np.abs(b["w"]).max())
    np.testing.assert_allclose(m["z"], b["z"], rtol=0, atol=1e-10 * np.abs(b["z"]).max())
    # distortionless in element space: a^H w = 1 (a = 1), the reference's w = y / sum(y)
    np.testing.assert_allclose(m["w"].sum(axis=-1), 1.0, atol=1e-10)


def test_beamspace_matrix_loads_in_element_space():
    """At d > 0 the matrix path loads S_L = S + d tr(S)/L I before the products, so
    R_bs = B S B^H + d tr(S)/L I_Nb (B B^H = I_Nb): the sliding-DFT path with the loading
    d' = d (tr R_hat / L) / (tr R_bs / Nb) per pixel (its convention is d' tr(R_bs)/Nb,
    CAP/getCaponAllC.pyx:376-385) must give the same P, w and z; with d itself it must not."""
    x, L, Y, Nb = _beamspace_case(22)
    d = 0.05
    K = x.shape[-1] - L + 1
    m = O.capon_beamspace_matrix(x, d, L, Y, Nb)
    ix, j = O.all_pixels(x.shape[0], x.shape[1], Y)
    rows = O.window_rows(x, ix, j, Y)
    Rhat = O.covariance_unnormalised(rows, L) / (K * (2 * Y + 1))
    v = O.sliding_dft(rows, L, Nb)                                     # beamspace subarrays
    tr_bs = np.einsum("btkr,btkr->b", v, v.conj()).real / (K * (2 * Y + 1))
    tr_es = np.trace(Rhat, axis1=-2, axis2=-1).real
    dprime = d * (tr_es / L) / (tr_bs / Nb)
    b = O.capon_beamspace(x, dprime, L, Y, Nb)
    np.testing.assert_allclose(m["P"], b["P"], rtol=1e-10)
    np.testing.assert_allclose(m["w"], b["w"], rtol=0, atol=1e-10 * np.abs(b["w"]).max())
    np.testing.assert_allclose(m["z"], b["z"], rtol=0, atol=1e-10 * np.abs(b["z"]).max())
    wrong = O.capon_beamspace(x, d, L, Y, Nb)
    assert np.max(np.abs(wrong["P"] - m["P"]) / m["P"]) > 1e-3
    # the materialised R is the element-space loaded S_L in the GPU layout
    SL = O.diagonal_load(Rhat * (K * (2 * Y + 1)), d, L)
    np.testing.assert_allclose(m["R"], np.conj(SL), rtol=0, atol=1e-12 * np.abs(SL).max())


def test_beamspace_matrix_white_covariance_closed_form():
    """Snapshots whose subarrays are orthogonal with equal energy give S = c I; then
    R_bs = (c + d c) I_Nb in the matrix path and P = (1 + d) c / (norm L) since
    a_bs = B 1 = sqrt(L) e_0: a closed form independent of both beamspace restatements."""
    L, Nb, Y = 8, 3, 0
    M = 2 * L - 1                        # K = L subarrays of one snapshot
    # x[m] = 1 at m = L-1 only: subarray k holds e_{L-1-k}, so S = sum_k e e^H = I
    x = np.zeros((1, 1, M), dtype=np.complex128)
    x[0, 0, L - 1] = 1.0
    d = 0.25
    m = O.capon_beamspace_matrix(x, d, L, Y, Nb)
    norm = L * (2 * Y + 1)
    # S_L = (1 + d) I, R_bs = (1 + d) I, a_bs^H R_bs^-1 a_bs = L / (1 + d)
    np.testing.assert_allclose(m["P"], (1 + d) / (L * norm), rtol=1e-12)
    # y_bs = sqrt(L)/(1+d) e_0, y = B^H y_bs = (1/(1+d)) 1, w = 1/L
    np.testing.assert_allclose(m["w"], np.full((1, L), 1.0 / L), atol=1e-14)


# ---------------------------------------------------------------- steering / rows

def test_steered_equals_rotated_dropin():
    rng = np.random.default_rng(10)
    M, L, Y, d = 12, 6, 1, 0.05
    x = rng.standard_normal((1, 6, M)) + 1j * rng.standard_normal((1, 6, M))
    sines = np.array([-0.4, 0.0, 0.3])
    s = O.capon_steered(x, d, L, Y, sines)
    for n, sn in enumerate(sines):
        xr = x * np.exp(-1j * np.pi * np.arange(M) * sn)
        o = O.capon(xr, d, L, Y)
        np.testing.assert_allclose(s["P"][:, n], o["P"], rtol=1e-11)
        np.testing.assert_allclose(s["z"][:, n], o["z"], rtol=1e-10)
        np.testing.assert_allclose(s["w"][:, n], o["w"], atol=1e-11)


def test_rows_uniform_equals_capon():
    rng = np.random.default_rng(11)
    Nz, M = 8, 10
    x = rng.standard_normal((2, Nz, M)) + 1j * rng.standard_normal((2, Nz, M))
    Lr = np.full(Nz, 4)
    Mr = np.full(Nz, M)
    off = np.zeros(Nz, int)
    dr = np.full(Nz, 0.1)
    r = O.capon_rows(x, dr, Lr, Mr, off, 1)
    o = O.capon(x, 0.1, 4, 1)
    np.testing.assert_allclose(r["P"], o["P"], rtol=1e-12)
    # a sub-aperture equals the drop-in call on the sliced cube
    Mr2 = np.full(Nz, 6)
    off2 = np.full(Nz, 2)
    r2 = O.capon_rows(x, dr, Lr, Mr2, off2, 1)
    o2 = O.capon(x[:, :, 2:8], 0.1, 4, 1)
    np.testing.assert_allclose(r2["P"], o2["P"], rtol=1e-12)


def test_condition_bound_holds():
    rng = np.random.default_rng(12)
    x = rng.standard_normal((1, 5, 16)) + 1j * rng.standard_normal((1, 5, 16))
    for d in (1e-1, 1e-3):
        rows = O.window_rows(x, np.array([0]), np.array([1]), 1)
        RL = O.diagonal_load(O.covariance_unnormalised(rows, 8), d, 8)[0]
        assert np.linalg.cond(RL) <= O.loaded_condition_bound(d, 8) * (1 + 1e-9)


# ------------------------------------------------------------------ reference options

def test_fb_average_equals_conjugate_reversed_data():
    """(R + J R^T J)/2 is the covariance of the data together with its conjugate-reversed
    copy (the backward snapshots J conj(u)): an independent path through the covariance."""
    rng = np.random.default_rng(21)
    rows = rng.standard_normal((3, 5, 14)) + 1j * rng.standard_normal((3, 5, 14))
    L = 6
    Rf = O.covariance_unnormalised(rows, L)
    Rb = O.covariance_unnormalised(np.conj(rows[..., ::-1]), L)
    np.testing.assert_allclose(O.fb_average(Rf), 0.5 * (Rf + Rb), rtol=0, atol=1e-12 * np.abs(Rf).max())
    F = O.fb_average(Rf)
    J = np.eye(L)[::-1]
    np.testing.assert_allclose(F, J @ np.swapaxes(F, -1, -2) @ J, atol=1e-12 * np.abs(F).max())   # persymmetric
    np.testing.assert_allclose(O.fb_average(F), F, atol=1e-12 * np.abs(F).max())                    # idempotent
    np.testing.assert_allclose(np.trace(F, axis1=-2, axis2=-1), np.trace(Rf, axis1=-2, axis2=-1), rtol=1e-12)


def test_capon_options_default_equals_capon():
    rng = np.random.default_rng(22)
    x = rng.standard_normal((2, 9, 12)) + 1j * rng.standard_normal((2, 9, 12))
    o = O.capon(x, 0.05, 5, 1)
    q = O.capon_options(x, 0.05, 5, 1)
    np.testing.assert_allclose(q["P"], o["P"], rtol=1e-12)
    np.testing.assert_allclose(q["z"], o["z"], rtol=1e-10)
    np.testing.assert_allclose(q["w"], o["w"], atol=1e-12)


def test_power_capon_constant_snapshot_closed_form():
    """A centre snapshot constant over the aperture gives w^H u_k = c (a^H w = 1) for every
    subarray, so the power-Capon amplitude is |c| / sqrt(K) whatever the weights."""
    rng = np.random.default_rng(23)
    M, L, Y = 16, 6, 1
    K = M - L + 1
    x = rng.standard_normal((1, 3, M)) + 1j * rng.standard_normal((1, 3, M))
    c = 0.7 - 1.3j
    x[0, 1, :] = c
    q = O.capon_options(x, 0.1, L, Y, z_mode=1, fb=True)
    np.testing.assert_allclose(q["z"][0].real, abs(c) / np.sqrt(K), rtol=1e-12)
    assert abs(q["z"][0].imag) == 0.0


def test_unloaded_power_closed_form_and_identity():
    """Snapshots sqrt(3c) e_t (L = M = 3, Yavg = 1) give R_hat = c I, w = 1/3:
    |w^H R_hat w| = c/3 and the loaded P = c(1+d)/3.  On random data the unloaded power
    equals P - (d tr(R_hat)/L) |w|^2."""
    c, d = 2.5, 0.2
    x = (np.sqrt(3 * c) * np.eye(3))[None].astype(np.complex128)
    q0 = O.capon_options(x, d, 3, 1)
    q1 = O.capon_options(x, d, 3, 1, p_mode=1)
    np.testing.assert_allclose(q1["P"], [c / 3], rtol=1e-12)
    np.testing.assert_allclose(q0["P"], [c * (1 + d) / 3], rtol=1e-12)
    rng = np.random.default_rng(24)
    x = rng.standard_normal((2, 7, 10)) + 1j * rng.standard_normal((2, 7, 10))
    L, Y = 4, 2
    q0 = O.capon_options(x, d, L, Y, fb=True)
    q1 = O.capon_options(x, d, L, Y, fb=True, p_mode=1)
    rows = O.window_rows(x, *O.all_pixels(2, 7, Y), Y)
    Rh = O.fb_average(O.covariance_unnormalised(rows, L) / ((10 - L + 1) * (2 * Y + 1)))
    tr = np.real(np.trace(Rh, axis1=-2, axis2=-1))
    np.testing.assert_allclose(q1["P"], q0["P"] - d * tr / L * np.sum(np.abs(q0["w"]) ** 2, axis=-1), rtol=1e-10)


def test_subspace_pins():
    """V = I reduces to element space; a unitary V leaves P unchanged (R_L = V R_hat_L V^H,
    a = V 1); the Butler rows reproduce the beamspace path (loading in beamspace)."""
    rng = np.random.default_rng(25)
    x = rng.standard_normal((2, 8, 14)) + 1j * rng.standard_normal((2, 8, 14))
    L, Y, d = 6, 1, 0.05
    e = O.capon_options(x, d, L, Y, fb=True)
    s = O.capon_subspace(x, d, L, Y, np.eye(L), fb=True)
    np.testing.assert_allclose(s["P"], e["P"], rtol=1e-12)
    np.testing.assert_allclose(s["z"], e["z"], rtol=1e-10)
    Q, _ = np.linalg.qr(rng.standard_normal((L, L)) + 1j * rng.standard_normal((L, L)))
    u = O.capon_subspace(x, d, L, Y, Q.conj().T)
    np.testing.assert_allclose(u["P"], O.capon(x, d, L, Y)["P"], rtol=1e-10)
    Nb = 3
    B = O.butler_matrix(Nb, L)
    b = O.capon_subspace(x, d, L, Y, B)
    r = O.capon_beamspace(x, d, L, Y, Nb)
    np.testing.assert_allclose(b["P"], r["P"], rtol=1e-10)
    np.testing.assert_allclose(b["z"], r["z"], rtol=1e-9)
    np.testing.assert_allclose(b["w"], r["w_bs"], atol=1e-10)
