"""CPU FP64 restatement of the reference's Capon/MVDR beamformer.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: it is
imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline
leg of ``bench.py``, and by nothing that ships.  The product path never calls it.

Each function restates one piece of the reference, cited file:line relative to
the reference root (``CAP/`` = framework/beamformer/capon/, ``RTC/`` =
CAP/RealTimeCapon/, ``LA/`` = framework/linalg/).  Everything is vectorised over
a batch of pixels with numpy in complex128; per-pixel loop order follows the
reference where it matters (Cholesky), elsewhere the arithmetic is the same
closed form evaluated in FP64.

Conventions (pixel p = (ix, j), centre snapshot c = j + Yavg, window c-Y..c+Y):
  u_{t,k}[l] = x[ix, t, k+l]                     k < K = M-L+1
  S          = sum_t sum_k u u^H                  (RTC/buildR_kernel.cu:96-137, unnormalised)
  R_hat      = S / (K (2Y+1))                     (CAP/getCaponC.pyx:240-242)
  R_L        = R_hat + d tr(R_hat)/L I            (CAP/getCaponC.pyx:258-262)
  y = R_L^-1 a,  w = y/(a^H y),  g_l = (1/K) sum_k x[ix,c,k+l],  z = w^H g,
  P = 1/Re(a^H y)                                 (CAP/getCaponC.pyx:274-289 with a = 1)

Pinning: every function here is checked in tests/test_oracle.py against
closed forms, brute force or the reference's own known-answer fixtures
(tests/golden/).  Functions without such a check are marked "parity unpinned".
"""
from __future__ import annotations

import numpy as np


# --------------------------------------------------------------------------------------
# Covariance
# --------------------------------------------------------------------------------------

def window_rows(x: np.ndarray, ix, j, Yavg: int) -> np.ndarray:
    """Snapshots x[ix, j .. j+2Y, :] for pixels (ix, j); returns [B][2Y+1][M] complex128."""
    ix = np.asarray(ix)
    j = np.asarray(j)
    t = j[:, None] + np.arange(2 * Yavg + 1)[None, :]
    return x[ix[:, None], t, :].astype(np.complex128)


def subarrays(rows: np.ndarray, L: int) -> np.ndarray:
    """u[b, t, k, l] = rows[b, t, k+l] (Hankel subarray matrix of each snapshot)."""
    M = rows.shape[-1]
    K = M - L + 1
    idx = np.arange(K)[:, None] + np.arange(L)[None, :]
    return rows[..., idx]


def covariance_unnormalised(rows: np.ndarray, L: int) -> np.ndarray:
    """S[b, r, c] = sum_t sum_k x_t[k+r] conj(x_t[k+c])  (RTC/buildR_kernel.cu:96-137;
    CAP/getCaponC.pyx:208-237 before the conj/normalisation at :240-242)."""
    u = subarrays(rows, L)
    return np.einsum("btkr,btkc->brc", u, u.conj(), optimize=True)


def diagonal_load(R: np.ndarray, d, L: int) -> np.ndarray:
    """R + d tr(R)/L I (CAP/getCaponC.pyx:258-262; RTC/buildR_kernel.cu:110-154 uses the
    real part of the trace)."""
    d = np.asarray(d, dtype=np.float64)
    tr = np.real(np.trace(R, axis1=-2, axis2=-1))
    out = R.copy()
    ii = np.arange(R.shape[-1])
    out[..., ii, ii] += (d * tr / L)[..., None]
    return out


def gpu_R_layout(S_L: np.ndarray) -> np.ndarray:
    """The reference GPU build_R output: memory[p][r][c] = conj(S_L[r][c]), i.e. the
    loaded unnormalised covariance in column-major order (RTC/buildR_kernel.cu:84-86,
    :156-157; pinned by CAP/tests/testRTC.pyx:180)."""
    return np.conj(S_L)


# --------------------------------------------------------------------------------------
# Cholesky solve (LA/matrixDecompc.pyx:19-50, LA/linearSolvec.pyx:259-346)
# --------------------------------------------------------------------------------------

def cholesky_upper(A: np.ndarray) -> np.ndarray:
    """U upper with U^H U = A, reading the upper triangle only (LA/matrixDecompc.pyx:19-50).
    Batched over leading axes; inner order follows the reference loop."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[-1]
    U = np.zeros_like(A)
    for i in range(n):
        s = np.sum(U[..., :i, i] * np.conj(U[..., :i, i]), axis=-1)
        U[..., i, i] = np.sqrt(A[..., i, i] - s)
        for j in range(i + 1, n):
            s = np.sum(np.conj(U[..., :i, i]) * U[..., :i, j], axis=-1)
            U[..., i, j] = (A[..., i, j] - s) / U[..., i, i]
    return U


def forward_solve(Lo: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Lower-triangular solve Lo x = b (LA/linearSolvec.pyx:312-327)."""
    n = Lo.shape[-1]
    x = np.zeros(np.broadcast_shapes(Lo.shape[:-1], b.shape), dtype=np.complex128)
    for i in range(n):
        x[..., i] = (b[..., i] - np.sum(Lo[..., i, :i] * x[..., :i], axis=-1)) / Lo[..., i, i]
    return x


def back_solve(U: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Upper-triangular solve U x = b (LA/linearSolvec.pyx:330-346)."""
    n = U.shape[-1]
    x = np.zeros(np.broadcast_shapes(U.shape[:-1], b.shape), dtype=np.complex128)
    for i in range(n - 1, -1, -1):
        x[..., i] = (b[..., i] - np.sum(U[..., i, i + 1:] * x[..., i + 1:], axis=-1)) / U[..., i, i]
    return x


def solve_cholesky(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """x = A^-1 b for Hermitian positive definite A (LA/linearSolvec.pyx:259-282)."""
    U = cholesky_upper(A)
    UT = np.conj(np.swapaxes(U, -1, -2))
    return back_solve(U, forward_solve(UT, b))


# --------------------------------------------------------------------------------------
# Element-space Capon (CAP/getCaponC.pyx:98-297; RTC/capon_kernel.cu:338-427)
# --------------------------------------------------------------------------------------

def subarray_mean(centre: np.ndarray, L: int) -> np.ndarray:
    """g_l = (1/K) sum_{k<K} x_c[k+l] (CAP/getCaponC.pyx:247-255; RTC/capon_kernel.cu:402-423)."""
    M = centre.shape[-1]
    K = M - L + 1
    idx = np.arange(K)[:, None] + np.arange(L)[None, :]
    return centre[..., idx].sum(axis=-2) / K


def all_pixels(Nx: int, Nz: int, Yavg: int):
    ix, j = np.meshgrid(np.arange(Nx), np.arange(Nz - 2 * Yavg), indexing="ij")
    return ix.ravel(), j.ravel()


def capon(x: np.ndarray, d: float, L: int, Yavg: int, pixels=None, ref_fp32_constants=False,
          want_R=False):
    """FP64 Capon for the given pixels (default: all, ix-major like the reference output).

    Returns dict with P [B], z [B], w [B, L], zPow_ref [B] (the reference CPU quirk: the
    quadratic form over the upper triangle only, CAP/getCaponC.pyx:285-289 with R zero
    below the diagonal from :180,235-242), y [B, L] = R_L^-1 1, and, if want_R, the
    GPU-layout loaded unnormalised R [B, L, L].
    ref_fp32_constants reproduces the reference's float32 regCoef/K_inv/norm_inv
    (CAP/getCaponC.pyx:79,102,129-132).
    """
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    S = covariance_unnormalised(rows, L)
    norm = K * (2 * Yavg + 1)
    if ref_fp32_constants:
        norm_inv = float(np.float32(1.0 / np.float32(norm)))
        k_inv = float(np.float32(1.0 / np.float32(K)))
        dd = float(np.float32(d))
    else:
        norm_inv, k_inv, dd = 1.0 / norm, 1.0 / K, float(d)
    Rh = S * norm_inv
    RL = diagonal_load(Rh, dd, L)
    a = np.ones(L, dtype=np.complex128)
    y = solve_cholesky(RL, np.broadcast_to(a, RL.shape[:-1]))
    aRia = np.sum(np.conj(a) * y, axis=-1)
    w = y / aRia[:, None]
    centre = rows[:, Yavg, :]
    M_ = centre.shape[-1]
    idx = np.arange(K)[:, None] + np.arange(L)[None, :]
    g = centre[..., idx].sum(axis=-2) * k_inv
    z = np.sum(np.conj(w) * g, axis=-1)
    Rup = np.triu(RL)
    zpow = np.einsum("bj,bij,bi->b", np.conj(w), Rup, w)
    out = dict(P=1.0 / np.real(aRia), z=z, w=w, zPow_ref=zpow, y=y, g=g)
    if want_R:
        out["R"] = gpu_R_layout(diagonal_load(S, float(d), L))
    return out


def fb_average(R: np.ndarray) -> np.ndarray:
    """Forward-backward averaging (R + J R^T J)/2, J the exchange matrix
    (CAP/getCaponAllC.pyx:357-358, applied to the normalised covariance before loading)."""
    return 0.5 * (R + np.flip(np.swapaxes(R, -1, -2), axis=(-2, -1)))


def capon_options(x: np.ndarray, d: float, L: int, Yavg: int, pixels=None, fb: bool = False,
                  z_mode: int = 0, p_mode: int = 0):
    """Capon with the reference's optional variants (FP64, a = 1):
      fb      forward-backward averaged covariance (fb_average) before the loading;
      z_mode  0: z = w^H g (RTC/capon_kernel.cu:338-427); 1: power Capon
              z = sqrt(sum_k |w^H u_{c,k}|^2) / K over the centre snapshot's subarrays
              (RTC/capon_kernel.cu:235-335, power_capon_kernel);
      p_mode  0: P = 1/(a^H R_L^-1 a); 1: P = |w^H R_hat w| with R_hat unloaded
              (CAP/getCaponMatlab_ref.m:109).
    Returns dict P, z, w."""
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    Rh = covariance_unnormalised(rows, L) / (K * (2 * Yavg + 1))
    if fb:
        Rh = fb_average(Rh)
    RL = diagonal_load(Rh, float(d), L)
    y = solve_cholesky(RL, np.ones(RL.shape[:-1], dtype=np.complex128))
    q = np.real(y.sum(axis=-1))
    w = y / q[:, None]
    centre = rows[:, Yavg, :]
    if z_mode == 0:
        z = np.sum(np.conj(w) * subarray_mean(centre, L), axis=-1)
    else:
        u = subarrays(centre[:, None, :], L)[:, 0]               # [B][K][L]
        s = np.einsum("bl,bkl->bk", np.conj(w), u)
        z = np.sqrt(np.sum(np.abs(s) ** 2, axis=-1)) / K + 0j
    if p_mode == 0:
        P = 1.0 / q
    else:
        P = np.abs(np.einsum("bi,bij,bj->b", np.conj(w), Rh, w))
    return dict(P=P, z=z, w=w)


def capon_subspace(x: np.ndarray, d: float, L: int, Yavg: int, V: np.ndarray, pixels=None,
                   fb: bool = False):
    """Capon in the subspace spanned by the rows of a user matrix V [Nb][L]
    (CAP/getCaponAllC.pyx:230,346-447 with DIAGONAL_LOADING_IN_BEAMSPACE, its default):
    R = V R_hat V^H (R_hat normalised, forward-backward averaged first when fb),
    loading d tr(R)/Nb, a = V 1, g = V g_es, y = R_L^-1 a, w = y/(a^H y), z = w^H g,
    P = 1/(a^H R_L^-1 a).  Returns dict P, z, w (Nb-dimensional)."""
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    V = np.asarray(V, dtype=np.complex128)
    Nb = V.shape[0]
    rows = window_rows(x, ix, j, Yavg)
    Rh = covariance_unnormalised(rows, L) / (K * (2 * Yavg + 1))
    if fb:
        Rh = fb_average(Rh)
    Rbs = V @ Rh @ V.conj().T
    RL = diagonal_load(Rbs, float(d), Nb)
    a = V @ np.ones(L, dtype=np.complex128)
    y = solve_cholesky(RL, np.broadcast_to(a, RL.shape[:-1]))
    q = np.sum(np.conj(a) * y, axis=-1)
    w = y / q[:, None]
    g = subarray_mean(rows[:, Yavg, :], L) @ V.T
    z = np.sum(np.conj(w) * g, axis=-1)
    return dict(P=1.0 / np.real(q), z=z, w=w)


def reference_gpu_outputs(x: np.ndarray, d: float, L: int, Yavg: int, Nb: int = 0):
    """What Capon::getCapon returns for the whole cube (RTC/Capon.cpp:73-463 with the
    default build: COPY_R_DEVICE_HOST on, SLIDING_BEAMSPACE off): z [Nx][Nz-2Y],
    w [Nx][Nz-2Y][L], R [Nx][Nz-2Y][L][L] (conj-stored, loaded, unnormalised)."""
    Nx, Nz, M = x.shape
    if Nb == 0:
        o = capon(x, d, L, Yavg, want_R=True)
        sh = (Nx, Nz - 2 * Yavg)
        return dict(z=o["z"].reshape(sh), w=o["w"].reshape(sh + (L,)),
                    R=o["R"].reshape(sh + (L, L)), P=o["P"].reshape(sh))
    o = capon_beamspace_matrix(x, d, L, Yavg, Nb)
    sh = (Nx, Nz - 2 * Yavg)
    return dict(z=o["z"].reshape(sh), w=o["w"].reshape(sh + (L,)),
                R=o["R"].reshape(sh + (L, L)), P=o["P"].reshape(sh))


# --------------------------------------------------------------------------------------
# Beamspace (RTC/capon_kernel.cu:47-99, RTC/sliding_beamspace_kernel.cu:26-164,
# RTC/Capon.cpp:285-402, CAP/getCaponAllC.pyx:346-447)
# --------------------------------------------------------------------------------------

def beam_index(Nb: int, L: int) -> np.ndarray:
    """Beams symmetric about broadside: beta(b) = b for b <= Nb/2, else b + L - 2(Nb/2) - 1.
    The even-Nb '+1' of RTC/capon_kernel.cu:58 never fires (`Nb&1 == 0` parses as
    `Nb & (1 == 0)`), so it is not applied."""
    b = np.arange(Nb)
    return np.where(b > Nb // 2, b + L - 2 * (Nb // 2) - 1, b)


def butler_matrix(Nb: int, L: int) -> np.ndarray:
    """B[b, l] = L^-1/2 exp(-j 2 pi beta(b) l / L) (RTC/capon_kernel.cu:47-74;
    LA/beamspace.py:18-34 for the first Nb rows)."""
    beta = beam_index(Nb, L)
    l = np.arange(L)
    return np.exp(-2j * np.pi * np.outer(beta, l) / L) / np.sqrt(L)


def sliding_dft(rows: np.ndarray, L: int, Nb: int) -> np.ndarray:
    """v[..., k, b] = sum_l B[b, l] x[k+l] by the sliding recursion
    S_{k+1} = (S_k + x[k+L] - x[k]) e^{+j 2 pi beta / L}, scaled by L^-1/2
    (RTC/sliding_beamspace_kernel.cu:26-81)."""
    M = rows.shape[-1]
    K = M - L + 1
    beta = beam_index(Nb, L)
    l = np.arange(L)
    tw0 = np.exp(-2j * np.pi * np.outer(l, beta) / L)          # [L][Nb]
    step = np.exp(2j * np.pi * beta / L)                         # [Nb]
    out = np.empty(rows.shape[:-1] + (K, Nb), dtype=np.complex128)
    s = rows[..., :L] @ tw0
    out[..., 0, :] = s
    for k in range(1, K):
        s = (s + rows[..., k + L - 1, None] - rows[..., k - 1, None]) * step
        out[..., k, :] = s
    return out / np.sqrt(L)


def capon_beamspace(x: np.ndarray, d: float, L: int, Yavg: int, Nb: int, pixels=None):
    """Beamspace Capon with loading in beamspace, d tr(R_bs)/Nb (CAP/getCaponAllC.pyx:
    367-385 with DIAGONAL_LOADING_IN_BEAMSPACE, the default at :19; also the reference's
    sliding GPU path RTC/Capon.cpp:202-227,399).
    R_bs = (1/(K(2Y+1))) sum v v^H, a_bs = B 1 = sqrt(L) e_0, w_bs = R_bs^-1 a_bs / (a^H ..),
    z = w_bs^H (B g), element-space weights w = B^H w_bs, P = 1/(a_bs^H R_bs^-1 a_bs)."""
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    v = sliding_dft(rows, L, Nb)                         # [B][T][K][Nb]
    Rbs = np.einsum("btkr,btkc->brc", v, v.conj(), optimize=True) / (K * (2 * Yavg + 1))
    RL = diagonal_load(Rbs, d, Nb)
    B = butler_matrix(Nb, L)
    abs_ = B @ np.ones(L)
    y = solve_cholesky(RL, np.broadcast_to(abs_, RL.shape[:-1]))
    q = np.sum(np.conj(abs_) * y, axis=-1)
    wbs = y / q[:, None]
    g = subarray_mean(rows[:, Yavg, :], L)
    gbs = g @ B.T
    z = np.sum(np.conj(wbs) * gbs, axis=-1)
    w = wbs @ np.conj(B)
    return dict(P=1.0 / np.real(q), z=z, w=w, w_bs=wbs)


def capon_beamspace_matrix(x: np.ndarray, d: float, L: int, Yavg: int, Nb: int, pixels=None):
    """The reference's default GPU beamspace path (RTC/Capon.cpp:285-402): element-space
    loaded unnormalised S_L, R_bs = B S_L B^H, a_bs = B 1, y_bs = R_bs^-1 a_bs,
    y = B^H y_bs, then the element-space output stage w = y/sum(y), z = w^H g.
    P (not produced by the reference) is 1/(a_bs^H R_bs'^-1 a_bs) for the normalised
    R_bs' = R_bs/(K(2Y+1)).  The reference computes the products with cublasCgemmBatched,
    whose only reference test is report-only (CAP/tests/testRTC.pyx:386-500); pinned in
    tests/test_oracle.py at Nb < L by the sliding-DFT restatement (d = 0: same P, w, z;
    d > 0: same with the per-pixel loading that element-space loading implies) and by the
    closed form for S = c I."""
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    S = covariance_unnormalised(rows, L)
    SL = diagonal_load(S, d, L)
    B = butler_matrix(Nb, L)
    Rbs = B @ SL @ np.conj(B.T)
    abs_ = B @ np.ones(L)
    ybs = solve_cholesky(Rbs, np.broadcast_to(abs_, Rbs.shape[:-1]))
    y = ybs @ np.conj(B)
    s = y.sum(axis=-1)
    w = y / s[:, None]
    g = subarray_mean(rows[:, Yavg, :], L)
    z = np.sum(np.conj(w) * g, axis=-1)
    q = np.sum(np.conj(abs_) * ybs, axis=-1)
    # a_bs^H (R_bs / norm)^-1 a_bs = norm * q
    P = 1.0 / (np.real(q) * (K * (2 * Yavg + 1)))
    return dict(P=P, z=z, w=w, R=gpu_R_layout(SL))


# --------------------------------------------------------------------------------------
# Extensions with no reference counterpart (steering, per-row sizes)
# --------------------------------------------------------------------------------------

def capon_steered(x: np.ndarray, d: float, L: int, Yavg: int, sines: np.ndarray, pixels=None):
    """One factorisation per pixel applied to many steering vectors a_theta[l] =
    exp(j pi l sin theta).  Equivalent to the drop-in call on the cube phase-rotated by
    exp(-j pi m sin theta) because R_theta = D R D^H leaves trace and loading unchanged.
    Returns P [B][A], w [B][A][L] (in the rotated, drop-in domain), z [B][A]."""
    Nx, Nz, M = x.shape
    K = M - L + 1
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    S = covariance_unnormalised(rows, L)
    RL = diagonal_load(S / (K * (2 * Yavg + 1)), d, L)
    l = np.arange(L)
    A = np.exp(1j * np.pi * np.outer(sines, l))                   # [A][L]
    U = cholesky_upper(RL)
    UT = np.conj(np.swapaxes(U, -1, -2))
    Bn = RL.shape[0]
    nA = len(sines)
    rhs = np.broadcast_to(A[None], (Bn, nA, L))
    zz = forward_solve(UT[:, None], rhs)
    y = back_solve(U[:, None], zz)
    q = np.sum(np.conj(A)[None] * y, axis=-1)
    w_nat = y / q[..., None]
    D = np.exp(-1j * np.pi * np.outer(sines, l))                  # rotation per angle
    w = w_nat * D[None]
    centre = rows[:, Yavg, :]
    m = np.arange(M)
    rot = centre[:, None, :] * np.exp(-1j * np.pi * np.outer(sines, m))[None]
    g = subarray_mean(rot, L)                                     # [B][A][L]
    z = np.sum(np.conj(w) * g, axis=-1)
    return dict(P=1.0 / np.real(q), w=w, z=z)


def capon_rows(x: np.ndarray, d_rows, L_rows, M_rows, off_rows, Yavg: int, pixels=None):
    """Depth-dependent aperture: the pixel centred on row c uses elements
    off(c) .. off(c)+M(c)-1 of every snapshot in its window, subarray length L(c) and
    loading d(c).  Returns P [B], z [B], w list-of-arrays (ragged)."""
    Nx, Nz, _ = x.shape
    if pixels is None:
        pixels = all_pixels(Nx, Nz, Yavg)
    ix, j = pixels
    ix = np.asarray(ix)
    j = np.asarray(j)
    c = j + Yavg
    P = np.zeros(len(ix))
    z = np.zeros(len(ix), dtype=np.complex128)
    w = [None] * len(ix)
    keys = np.stack([L_rows[c], M_rows[c], off_rows[c]], axis=1)
    for key in np.unique(keys, axis=0):
        L, M, off = (int(v) for v in key)
        sel = np.where((keys == key).all(axis=1))[0]
        dsel = np.asarray(d_rows)[c[sel]].astype(np.float64)
        sub = x[:, :, off:off + M]
        o = _capon_batch(sub, dsel, L, Yavg, (ix[sel], j[sel]))
        P[sel] = o["P"]
        z[sel] = o["z"]
        for n, s in enumerate(sel):
            w[s] = o["w"][n]
    return dict(P=P, z=z, w=w)


def _capon_batch(x, d_vec, L, Yavg, pixels):
    Nx, Nz, M = x.shape
    K = M - L + 1
    ix, j = pixels
    rows = window_rows(x, ix, j, Yavg)
    S = covariance_unnormalised(rows, L)
    RL = diagonal_load(S / (K * (2 * Yavg + 1)), d_vec, L)
    y = solve_cholesky(RL, np.ones(RL.shape[:-1], dtype=np.complex128))
    q = y.sum(axis=-1)
    w = y / q[:, None]
    g = subarray_mean(rows[:, Yavg, :], L)
    return dict(P=1.0 / np.real(q), z=np.sum(np.conj(w) * g, axis=-1), w=w)


def loaded_condition_bound(d, L) -> np.ndarray:
    """Upper bound on cond(R_L): lambda_max <= tr(R_hat), lambda_min >= d tr(R_hat)/L,
    so cond <= 1 + L/d."""
    return 1.0 + np.asarray(L, dtype=np.float64) / np.asarray(d, dtype=np.float64)
