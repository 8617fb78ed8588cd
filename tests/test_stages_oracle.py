"""Pins of the delay-stage (TDIB) and window-selection (LCA) restatements (CPU)."""
import numpy as np
import pytest

from oracle import stages_oracle as S


def _geometry(rng, n_h=8, n_x=5, n_y=40):
    x_r = np.linspace(0.0, 1.2, n_h)
    D = 0.3 + 0.02 * rng.standard_normal(n_h)
    xarr = np.repeat(np.linspace(-2.0, 2.0, n_x)[:, None], n_y, axis=1)
    yarr = np.repeat(np.linspace(20.0, 60.0, n_y)[None, :], n_x, axis=0)
    return x_r, D, xarr, yarr


def test_tdib_interpolation_matches_scipy_interp1d():
    """The reference interpolates with scipy.interpolate.interp1d(kind='linear',
    bounds_error=False) on real and imaginary parts (tdibC.pyx:51-54)."""
    from scipy import interpolate
    rng = np.random.default_rng(1)
    c, fc = 1500.0, 1.0e5
    x_r, D, xarr, yarr = _geometry(rng)
    mtaxe = np.linspace(0.02, 0.09, 900)
    mf = rng.standard_normal((8, 900)) + 1j * rng.standard_normal((8, 900))
    cube = S.tdib_cube(mf, mtaxe, x_r, D, c, fc, xarr, yarr)
    xx, zz = np.mean(x_r) / 2, np.mean(D) / 2
    for n in (0, 5):
        r_t = np.sqrt((xx + xarr) ** 2 + yarr ** 2 + zz ** 2)
        r_r = np.sqrt((xx + xarr - x_r[n]) ** 2 + yarr ** 2 + (zz - D[n]) ** 2)
        t = (r_t + r_r) / c
        ref = (interpolate.interp1d(mtaxe, mf[n].real, kind="linear", bounds_error=False)(t)
               + 1j * interpolate.interp1d(mtaxe, mf[n].imag, kind="linear", bounds_error=False)(t))
        ref = ref * np.exp(2j * np.pi * fc * t)
        np.testing.assert_array_equal(np.isnan(cube[..., n]), np.isnan(ref))
        ok = ~np.isnan(ref)
        np.testing.assert_allclose(cube[..., n][ok], ref[ok], rtol=1e-12, atol=1e-12)


def test_tdib_linear_signal_is_exact_and_geometry_scalar():
    """Data linear in time is interpolated exactly; the travel time of one pixel re-derived
    with scalar arithmetic; the image is the hydrophone mean of the cube, transposed."""
    rng = np.random.default_rng(2)
    c, fc = 1480.0, 9.0e4
    x_r, D, xarr, yarr = _geometry(rng, n_h=4, n_x=3, n_y=11)
    mtaxe = np.linspace(0.0, 0.2, 4001)
    a, b = 0.3 - 0.2j, 5.0 + 1.0j
    mf = np.tile(a + b * mtaxe, (4, 1))
    cube = S.tdib_cube(mf, mtaxe, x_r, D, c, fc, xarr, yarr)
    n, ix, iy = 2, 1, 7
    xx, zz = sum(x_r) / len(x_r) / 2, sum(D) / len(D) / 2
    X, Yp = float(xarr[ix, iy]), float(yarr[ix, iy])
    t = (((xx + X) ** 2 + Yp ** 2 + zz ** 2) ** 0.5 + ((xx + X - x_r[n]) ** 2 + Yp ** 2 + (zz - D[n]) ** 2) ** 0.5) / c
    expect = (a + b * t) * complex(np.cos(2 * np.pi * fc * t), np.sin(2 * np.pi * fc * t))
    assert abs(cube[ix, iy, n] - expect) <= 1e-9 * abs(expect)
    img = S.tdib(mf, mtaxe, x_r, D, c, fc, xarr, yarr)
    np.testing.assert_allclose(img, cube.mean(axis=-1).T, rtol=1e-13)


def test_tdib_outside_axis_is_nan():
    rng = np.random.default_rng(3)
    x_r, D, xarr, yarr = _geometry(rng, n_h=2, n_x=2, n_y=5)
    mtaxe = np.linspace(0.0, 1e-3, 10)          # far too short for 20-60 m
    cube = S.tdib_cube(np.ones((2, 10)), mtaxe, x_r, D, 1500.0, 1e5, xarr, yarr)
    assert np.isnan(cube).all()


def test_lca_brute_force_and_special_cases():
    rng = np.random.default_rng(4)
    Nx, Ny, Nw = 6, 9, 4
    Z = rng.standard_normal((Nx, Ny, Nw)) + 1j * rng.standard_normal((Nx, Ny, Nw))
    wrf = rng.uniform(0.2, 1.0, (3, 3))
    img, win, _ = S.lca(Z, wrf)
    for x in range(Nx):
        for y in range(Ny):
            if 1 <= x < Nx - 1 and 1 <= y < Ny - 1:
                s = [sum(wrf[yy, xx] * abs(Z[x + xx - 1, y + yy - 1, w]) ** 2 for yy in range(3) for xx in range(3))
                     for w in range(Nw)]
                w = int(np.argmin(s))
                assert win[x, y] == w and img[x, y] == Z[x, y, w]
            else:
                assert win[x, y] == 0 and img[x, y] == 0
    # 1 x 1 filter: the per-pixel argmin of |Z| (lcacx.pyx:119-125)
    img1, win1, _ = S.lca(Z, np.ones((1, 1)))
    np.testing.assert_array_equal(win1, np.abs(Z).argmin(axis=2))
    # a window that is smallest everywhere is always chosen
    Z2 = Z.copy()
    Z2[..., 2] *= 1e-3
    _, win2, _ = S.lca(Z2, wrf)
    assert (win2[1:-1, 1:-1] == 2).all()
