"""CPU FP64 restatements of the stages either side of the beamformer call.

TEST INFRASTRUCTURE ONLY (see capon_oracle.py): imported by ``tests/`` and nothing that
ships.  Paths are relative to the reference root.

  tdib  the time-domain interpolation beamformer, framework/beamformer/delay/tdibC.pyx:9-67
  lca   the low-complexity adaptive window selection, framework/beamformer/lca/lcacx.pyx:27-217

  lca_reference  the same with the reference's own output, quirks included

Pinning (tests/test_stages_oracle.py): tdib's interpolation against scipy's
interp1d(kind="linear", bounds_error=False) -- the routine the reference calls -- and
against data linear in time (exact), the geometry against a scalar re-derivation, the image
as the hydrophone mean of the cube; lca against brute-force loops, the 1 x 1 filter against
the per-pixel argmin of |Z| (lcacx.pyx:119-125), and a window that is smallest everywhere;
lca_reference against lca on data whose sums are exact (rows 1+Ym .. Ny-Ym-1 of the
all-ones range filter, the transposed image for the general filter on square images),
against hand-worked cases where the reference's summation order decides the window, and
against the rows the reference leaves at zero.
"""
from __future__ import annotations

import numpy as np


def tdib_cube(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr):
    """Delayed, carrier-mixed channels v[nx][ny][n] = interp(mfdata[n], t) exp(j 2 pi fc t),
    t = (r_t + r_r)/c (tdibC.pyx:28-62); NaN where t is outside mtaxe (interp1d with
    bounds_error=False)."""
    mfdata = np.asarray(mfdata, dtype=np.complex128)
    mtaxe = np.asarray(mtaxe, dtype=np.float64)
    x_r = np.asarray(x_r, dtype=np.float64)
    D = np.asarray(D, dtype=np.float64)
    xarr = np.asarray(xarr, dtype=np.float64)
    yarr = np.asarray(yarr, dtype=np.float64)
    n_h = mfdata.shape[0]
    xx = np.mean(x_r) / 2
    zz = np.mean(D) / 2
    r_t = np.sqrt((xx + xarr) ** 2 + yarr ** 2 + zz ** 2)                       # [n_x][n_y]
    out = np.empty(xarr.shape + (n_h,), dtype=np.complex128)
    for n in range(n_h):
        r_r = np.sqrt((xx + xarr - x_r[n]) ** 2 + yarr ** 2 + (zz - D[n]) ** 2)
        t = (r_t + r_r) / c
        re = np.interp(t, mtaxe, mfdata[n].real, left=np.nan, right=np.nan)
        im = np.interp(t, mtaxe, mfdata[n].imag, left=np.nan, right=np.nan)
        out[..., n] = (re + 1j * im) * np.exp(2j * np.pi * fc * t)
    return out


def tdib(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr):
    """The reference's return value: image.T / n_h, i.e. [n_y][n_x] (tdibC.pyx:64-67)."""
    cube = tdib_cube(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr)
    return cube.sum(axis=-1).T / cube.shape[-1]


def lca(Z, wrf):
    """Window selection (lcacx.pyx:27-217 with the intended [x][y] neighbourhood):
    s_w = sum wrf[yy][xx] |Z[x+xx-Xm][y+yy-Ym][w]|^2 over the filter support, first argmin,
    img = Z at that window; pixels whose support leaves the image get 0 and window 0."""
    Z = np.asarray(Z, dtype=np.complex128)
    wrf = np.asarray(wrf, dtype=np.float64)
    Nx, Ny, Nw = Z.shape
    wy, wx = wrf.shape
    Ym, Xm = wy // 2, wx // 2
    p = Z.real ** 2 + Z.imag ** 2
    img = np.zeros((Nx, Ny), dtype=np.complex128)
    win = np.zeros((Nx, Ny), dtype=np.int64)
    s = np.zeros((Nx - 2 * Xm, Ny - 2 * Ym, Nw))
    for yy in range(wy):
        for xx in range(wx):
            s += wrf[yy, xx] * p[xx:xx + Nx - 2 * Xm, yy:yy + Ny - 2 * Ym, :]
    w = np.argmin(s, axis=-1)                                                   # first minimum
    xs, ys = np.meshgrid(np.arange(Xm, Nx - Xm), np.arange(Ym, Ny - Ym), indexing="ij")
    img[Xm:Nx - Xm, Ym:Ny - Ym] = Z[xs, ys, w]
    win[Xm:Nx - Xm, Ym:Ny - Ym] = w
    return img, win, s


def _first_argmin(s):
    """Index of the first smallest entry along the last axis by the reference's test
    ``w_sum[w] < w_sum_min or w == 0`` (lcacx.pyx:161,177,212): NaNs never win."""
    best = s[..., 0].copy()
    bw = np.zeros(s.shape[:-1], dtype=np.int64)
    for w in range(1, s.shape[-1]):
        m = s[..., w] < best
        best = np.where(m, s[..., w], best)
        bw = np.where(m, w, bw)
    return bw


def lca_reference(Z, wrf):
    """lcacx.pyx:27-217 as it computes, returning (img [Nx][Ny], win [Nx][Ny]).

    Branches, in the reference's order:
      1 x 1 filter with a nonzero weight (:122-127): argmin over w of |Z| at every pixel
        (numpy's abs and argmin, as the reference).
      all-ones range filter, X_AVG = 1 (:94-100, :139-182): per beam, FP64 sliding sums
        s_w = (sum of the first Y_AVG rows), then s_w += p[y+Ym] - p[y-Ym-1] for
        y = 1+Ym .. Ny-Ym-1 (the difference rounded first), p = re^2 + im^2; those rows get
        the first minimum.  The first segment's window is written to row Y_AVG (not Y_MID)
        and always as window 0 (:151-164: w_sum_min is reset to 0 for every w), and the
        sliding loop overwrites it unless Y_AVG >= Ny - Ym; row Ym and the edge rows stay 0.
      any other filter (:185-217): s_w = sum_yy sum_xx A[w, y+yy-Ym, x+xx-Xm] wrf[yy, xx]
        in that order from 0, where A is the |Z|^2 array allocated [Nw][Nx][Ny] and filled
        as A[w, x, y] (:188-191) but read with its second index from y and its third from x
        (:206): flat index w Nx Ny + (y+yy-Ym) Ny + (x+xx-Xm).  For square images that is
        the transposed neighbourhood.  Pixels y in [Ym, Ny-Ym), x in [Xm, Nx-Xm) get the
        first minimum, the rest 0.
    Shapes where the reference reads or writes outside its arrays raise ValueError: the
    all-ones branch with Ny <= Y_AVG (its first-segment write), the general branch with
    Ny > Nx (its last window's reads)."""
    Z = np.asarray(Z, dtype=np.complex128)
    wrf = np.atleast_2d(np.asarray(wrf, dtype=np.float64))
    Nx, Ny, Nw = Z.shape
    YA, XA = wrf.shape
    Ym, Xm = (YA - 1) // 2, (XA - 1) // 2
    if YA % 2 == 0 or XA % 2 == 0:
        raise ValueError("the wrf filter must have odd sizes")
    img = np.zeros((Nx, Ny), dtype=np.complex128)
    win = np.zeros((Nx, Ny), dtype=np.int64)
    p = Z.real * Z.real + Z.imag * Z.imag                       # [Nx][Ny][Nw]
    if YA == 1 and XA == 1 and wrf[0, 0] != 0:
        win = np.abs(Z).argmin(2)
        img = np.take_along_axis(Z, win[..., None], axis=2)[..., 0]
        return img, win
    if XA == 1 and bool(np.all(wrf[:, 0] == 1)):
        if Ny <= YA:
            raise ValueError("the reference writes row Y_AVG outside the image when Ny <= Y_AVG")
        s = np.zeros((Nx, Nw))
        for y in range(YA):
            s = s + p[:, y, :]
        if YA >= Ny - Ym:                                       # the first segment's write survives
            win[:, YA] = 0
            img[:, YA] = Z[:, YA, 0]
        for y in range(1 + Ym, Ny - Ym):
            s = s + (p[:, y + Ym, :] - p[:, y - Ym - 1, :])
            bw = _first_argmin(s)
            win[:, y] = bw
            img[:, y] = Z[np.arange(Nx), y, bw]
        return img, win
    ys, xs = np.arange(Ym, Ny - Ym), np.arange(Xm, Nx - Xm)
    if len(ys) == 0 or len(xs) == 0:
        return img, win
    if Ny > Nx:
        raise ValueError("the reference's general branch reads past its |Z|^2 array when Ny > Nx")
    A = p.transpose(2, 0, 1).ravel()                            # A[w][x][y] flattened
    X, Yy = np.meshgrid(xs, ys, indexing="ij")                  # pixels [nx][ny]
    wv = np.arange(Nw)
    s = np.zeros(X.shape + (Nw,))
    for yy in range(YA):
        for xx in range(XA):
            f = wv * (Nx * Ny) + ((Yy + yy - Ym) * Ny + (X + xx - Xm))[..., None]
            s = s + A[f] * wrf[yy, xx]
    bw = _first_argmin(s)
    win[X, Yy] = bw
    img[X, Yy] = Z[X, Yy, bw]
    return img, win
