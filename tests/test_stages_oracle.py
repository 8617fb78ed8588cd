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
    uniform = np.linspace(0.02, 0.09, 900)
    jittered = 0.02 + np.cumsum(rng.uniform(0.3, 1.7, 900)) * (0.07 / 900)   # any sorted axis
    for mtaxe in (uniform, jittered):
        _check_interp1d(rng, c, fc, x_r, D, xarr, yarr, mtaxe)


def _check_interp1d(rng, c, fc, x_r, D, xarr, yarr, mtaxe):
    from scipy import interpolate
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


# ---------------------------------------------------------------- lca as the reference computes it

def _int_cube(rng, Nx, Ny, Nw):
    # small integers: every |Z|^2, product and sum below is exact in FP64, so summation
    # order cannot change a window and the reference's output is fixed by its indexing
    return (rng.integers(-6, 7, (Nx, Ny, Nw)) + 1j * rng.integers(-6, 7, (Nx, Ny, Nw))).astype(np.complex128)


def test_lca_reference_all_ones_range_filter_rows():
    """All-ones range filter (lcacx.pyx:139-182): rows 1+Ym .. Ny-Ym-1 equal the centred
    selection; row Ym (whose window the first segment computes) and the edge rows stay 0;
    the first segment's write to row Y_AVG is overwritten by the sliding loop."""
    rng = np.random.default_rng(41)
    Nx, Ny, Nw = 5, 17, 6
    Z = _int_cube(rng, Nx, Ny, Nw)
    for YA in (3, 5):
        Ym = YA // 2
        wrf = np.ones((YA, 1))
        img, win = S.lca_reference(Z, wrf)
        ci, cw, _ = S.lca(Z, wrf)
        rows = np.arange(1 + Ym, Ny - Ym)
        np.testing.assert_array_equal(win[:, rows], cw[:, rows])
        np.testing.assert_array_equal(img[:, rows], ci[:, rows])
        dead = np.setdiff1d(np.arange(Ny), rows)
        assert Ym in dead
        assert (win[:, dead] == 0).all() and (img[:, dead] == 0).all()


def test_lca_reference_first_segment_row_survives_on_short_images():
    """Ny - Ym <= Y_AVG < Ny: the sliding loop ends before row Y_AVG, which keeps the first
    segment's write, window 0 and Z[x, Y_AVG, 0] (lcacx.pyx:158-164, w_sum_min reset per w)."""
    rng = np.random.default_rng(42)
    Z = _int_cube(rng, 3, 4, 5)
    Z[..., 0] *= 10                                        # window 0 is not the minimum
    img, win = S.lca_reference(Z, np.ones((3, 1)))
    ci, cw, _ = S.lca(Z, np.ones((3, 1)))
    np.testing.assert_array_equal(win[:, 2], cw[:, 2])     # the one sliding row
    assert (win[:, 3] == 0).all()
    np.testing.assert_array_equal(img[:, 3], Z[:, 3, 0])
    assert (win[:, [0, 1]] == 0).all() and (img[:, [0, 1]] == 0).all()
    with pytest.raises(ValueError):
        S.lca_reference(Z[:, :3], np.ones((3, 1)))         # Ny <= Y_AVG


def test_lca_reference_sliding_sum_rounding_decides():
    """The FP64 sliding recursion, not the window sum, picks the window: |Z|^2 = 2^53 in
    row 0 of window 0 absorbs the 1s that follow (2^53 + 1 rounds to 2^53), so after it
    leaves, window 0's sum at row 2 is 1 although rows 1..3 hold 3; window 1 (exact sum 2)
    loses to it.  A direct sum would choose window 1."""
    Z = np.zeros((1, 8, 2), dtype=np.complex128)
    Z[0, 0, 0] = 2.0 ** 26 + 1j * 2.0 ** 26                 # |Z|^2 = 2^53
    Z[0, 1:, 0] = 1.0
    Z[0, 1:3, 1] = 1.0
    img, win = S.lca_reference(Z, np.ones((3, 1)))
    _, cw, _ = S.lca(Z, np.ones((3, 1)))
    assert cw[0, 2] == 1 and win[0, 2] == 0
    assert img[0, 2] == Z[0, 2, 0]


def test_lca_reference_general_filter_square_is_transposed():
    """Any other filter (lcacx.pyx:185-217) on a square image reads the |Z|^2 array with x
    and y swapped: its windows are the centred selection of the transposed image, and img
    takes Z at the unswapped pixel."""
    rng = np.random.default_rng(43)
    N, Nw = 9, 5
    Z = _int_cube(rng, N, N, Nw)
    for wrf in (rng.integers(1, 4, (3, 5)).astype(float), rng.integers(1, 4, (5, 3)).astype(float),
                np.ones((1, 3)), np.array([[1.0, 2.0, 1.0], [2.0, 4.0, 2.0], [1.0, 2.0, 1.0]])):
        img, win = S.lca_reference(Z, wrf)
        _, cw, _ = S.lca(Z.transpose(1, 0, 2), wrf)
        Ym, Xm = wrf.shape[0] // 2, wrf.shape[1] // 2
        inner = (slice(Xm, N - Xm), slice(Ym, N - Ym))
        np.testing.assert_array_equal(win[inner], cw[inner])
        xs, ys = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
        np.testing.assert_array_equal(img[inner], Z[xs, ys, win][inner])
        mask = np.ones((N, N), bool)
        mask[inner] = False
        assert (win[mask] == 0).all() and (img[mask] == 0).all()


def test_lca_reference_general_filter_order_and_shapes():
    """The general branch sums in the order yy, then xx from 0: with 2^53 first, the 1s that
    follow are lost, so windows 0 ([2^53, 1, 1]) and 1 ([2^53, 0, 1]) tie and the first
    wins, where the exact sums would choose window 1.  Wider-than-tall images (Ny < Nx)
    read on through the flat array; taller ones (Ny > Nx) are refused."""
    Z = np.zeros((3, 3, 2), dtype=np.complex128)
    big = 2.0 ** 26 + 1j * 2.0 ** 26
    Z[1, :, 0] = [big, 1, 1]       # flat row y = 1 of window 0 is Z[1, 0..2, 0]
    Z[1, :, 1] = [big, 0, 1]
    img, win = S.lca_reference(Z, np.ones((1, 3)))
    assert win[1, 1] == 0
    # Ny < Nx: the pixel reads A[w Nx Ny + (y+yy-Ym) Ny + (x+xx-Xm)] with the x offset
    # running past the row end into the next row
    rng = np.random.default_rng(44)
    Z = _int_cube(rng, 7, 4, 3)
    wrf = np.array([[1.0, 2.0, 3.0]])
    img, win = S.lca_reference(Z, wrf)
    A = (Z.real ** 2 + Z.imag ** 2).transpose(2, 0, 1).ravel()
    for x in range(1, 6):
        for y in range(4):
            s = [sum(A[w * 28 + y * 4 + x + xx - 1] * wrf[0, xx] for xx in range(3)) for w in range(3)]
            assert win[x, y] == int(np.argmin(s))
    with pytest.raises(ValueError):
        S.lca_reference(_int_cube(rng, 4, 7, 3), wrf)


def test_lca_reference_one_by_one_filter():
    rng = np.random.default_rng(45)
    Z = rng.standard_normal((4, 6, 5)) + 1j * rng.standard_normal((4, 6, 5))
    img, win = S.lca_reference(Z, np.array([[0.5]]))
    np.testing.assert_array_equal(win, np.abs(Z).argmin(axis=2))
    np.testing.assert_array_equal(img, np.take_along_axis(Z, win[..., None], 2)[..., 0])
    # a zero 1 x 1 weight falls to the general branch: every sum is 0, window 0 everywhere
    # (on an image no taller than wide, which that branch needs)
    Zw = Z.transpose(1, 0, 2)
    img0, win0 = S.lca_reference(Zw, np.array([[0.0]]))
    assert (win0 == 0).all()
    np.testing.assert_array_equal(img0, Zw[..., 0])
