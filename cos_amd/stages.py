"""Python mirrors of the stages either side of the beamformer call, on torch tensors.

``tdib`` replaces framework/beamformer/delay/tdibC.pyx:9-67 (same argument order and
meaning; returns the image [n_y][n_x]), ``tdib_cube`` also returns the delayed channels in
the Capon input layout [n_x][n_y][n_h]; ``lca`` replaces
framework/beamformer/lca/lcacx.pyx:27-217 (returns the selected image and window index).
Both run the HIP kernels of libcos_capon.so and raise if it is not built.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, lib


def _dev(a, dtype):
    import torch
    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda", dtype=dtype)
    else:
        t = torch.as_tensor(np.asarray(a), dtype=dtype, device="cuda")
    return t.contiguous()


def _axis(mtaxe):
    """(t0, dt, None) for a uniform axis (cos_tdib_run), else (0, 0, device float64 axis) for
    cos_tdib_run_axis; the axis must increase strictly."""
    import torch
    t = np.asarray(mtaxe.detach().cpu() if hasattr(mtaxe, "detach") else mtaxe, dtype=np.float64).ravel()
    if t.size == 1:
        return float(t[0]), 1.0, None
    if not np.all(np.diff(t) > 0):
        raise ValueError("mtaxe must increase strictly")
    dt = (t[-1] - t[0]) / (t.size - 1)
    if np.max(np.abs(np.diff(t) - dt)) <= 1e-9 * dt:
        return float(t[0]), float(dt), None
    return 0.0, 0.0, torch.as_tensor(t, dtype=torch.float64, device="cuda")


def _run_tdib(mf, n_h, n_t, mtaxe, xr, dd, c, fc, xa, ya, n_x, n_y, fill, cube, image, stream):
    t0, dt, ax = _axis(mtaxe)
    args = (xr.data_ptr(), dd.data_ptr(), float(c), float(fc), xa.data_ptr(), ya.data_ptr(), n_x, n_y, float(fill),
            cube.data_ptr() if cube is not None else None, image.data_ptr() if image is not None else None, stream)
    if ax is None:
        check(lib().cos_tdib_run(mf.data_ptr(), n_h, n_t, t0, dt, *args))
    else:
        if ax.numel() != n_t:
            raise ValueError("mtaxe must have n_t samples")
        check(lib().cos_tdib_run_axis(mf.data_ptr(), n_h, n_t, ax.data_ptr(), *args))


def tdib_cube(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr, *, fill=float("nan"), want_image=True, stream=None):
    """Delay stage: returns (cube [n_x][n_y][n_h] complex64, image [n_y][n_x] complex64 or None).
    mfdata [n_h][n_t]; mtaxe [n_t] strictly increasing (uniform or not); x_r, D [n_h];
    xarr, yarr [n_x][n_y] (xarr may be [n_x], broadcast along the beam as in the reference)."""
    import torch
    mf = _dev(mfdata, torch.complex64)
    if mf.dim() != 2:
        raise ValueError("mfdata must be [n_h][n_t] (several banks are not supported)")
    n_h, n_t = mf.shape
    ya = _dev(yarr, torch.float64)
    xa = _dev(xarr, torch.float64)
    if xa.dim() == 1:
        xa = xa[:, None].expand_as(ya).contiguous()
    n_x, n_y = ya.shape
    xr = _dev(x_r, torch.float64)
    dd = _dev(D, torch.float64)
    cube = torch.empty((n_x, n_y, n_h), dtype=torch.complex64, device="cuda")
    image = torch.empty((n_y, n_x), dtype=torch.complex64, device="cuda") if want_image else None
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    _run_tdib(mf, n_h, n_t, mtaxe, xr, dd, c, fc, xa, ya, n_x, n_y, fill, cube, image, stream)
    return cube, image


def tdib(mfdata, mtaxe, x_r, D, c, fc, xarr, yarr, *, fill=float("nan"), stream=None):
    """The reference's tdib(): the delay-and-sum image [n_y][n_x] (complex64 CUDA tensor)."""
    import torch
    mf = _dev(mfdata, torch.complex64)
    n_h, n_t = mf.shape
    ya = _dev(yarr, torch.float64)
    xa = _dev(xarr, torch.float64)
    if xa.dim() == 1:
        xa = xa[:, None].expand_as(ya).contiguous()
    n_x, n_y = ya.shape
    xr = _dev(x_r, torch.float64)
    dd = _dev(D, torch.float64)
    image = torch.empty((n_y, n_x), dtype=torch.complex64, device="cuda")
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    _run_tdib(mf, n_h, n_t, mtaxe, xr, dd, c, fc, xa, ya, n_x, n_y, fill, None, image, stream)
    return image


def lca(Z, wrf, *, mode="reference", stream=None):
    """Window selection: Z [Nx][Ny][Nw] complex, wrf [wy][wx] weights (odd sizes).
    mode "reference": lcacx.pyx's own output (its branches, its row-Y_MID gap and swapped
    neighbourhood, FP64 sums in its order; refuses the shapes where it reads or writes
    outside its arrays); "centred": the intended neighbourhood at every interior pixel.
    Returns (img [Nx][Ny] complex64, win [Nx][Ny] int32) CUDA tensors."""
    import torch
    z = _dev(Z, torch.complex64)
    w = _dev(wrf, torch.float32)          # a CUDA tensor is used where it is
    if w.dim() < 2:                       # as numpy.atleast_2d: a vector is one row
        w = w.reshape(1, -1).contiguous()
    Nx, Ny, Nw = z.shape
    wy, wx = w.shape
    m = {"reference": 0, "centred": 1}[mode]
    img = torch.empty((Nx, Ny), dtype=torch.complex64, device="cuda")
    win = torch.empty((Nx, Ny), dtype=torch.int32, device="cuda")
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    check(lib().cos_lca_run(z.data_ptr(), Nx, Ny, Nw, w.data_ptr(), wy, wx, m, img.data_ptr(), win.data_ptr(),
                            stream))
    return img, win
