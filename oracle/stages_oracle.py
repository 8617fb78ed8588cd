"""CPU FP64 restatements of the stages either side of the beamformer call.

TEST INFRASTRUCTURE ONLY (see capon_oracle.py): imported by ``tests/`` and nothing that
ships.  Paths are relative to the reference root.

  tdib  the time-domain interpolation beamformer, framework/beamformer/delay/tdibC.pyx:9-67
  lca   the low-complexity adaptive window selection, framework/beamformer/lca/lcacx.pyx:27-217

Pinning (tests/test_stages_oracle.py): tdib's interpolation against scipy's
interp1d(kind="linear", bounds_error=False) -- the routine the reference calls -- and
against data linear in time (exact), the geometry against a scalar re-derivation, the image
as the hydrophone mean of the cube; lca against brute-force loops, the 1 x 1 filter against
the per-pixel argmin of |Z| (lcacx.pyx:119-125), and a window that is smallest everywhere.
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
    p = np.abs(Z) ** 2
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
