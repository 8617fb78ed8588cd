"""Differentiable Capon power image for PyTorch.

``capon_power(x, d, L, Yavg)`` returns P (float32, one value per pixel) for a complex64
CUDA cube and is differentiable with respect to the cube (complex: torch's
d/dRe + j d/dIm convention) and to the scalar loading d.  The backward pass is the
library's hand-written adjoint kernel (cos_capon_backward): it recomputes the covariance
and its Cholesky factor rather than storing them.  Phase corrections, frames and losses
are ordinary torch code around it, e.g.

    xs = frames * torch.exp(1j * phi)            # [F][Nx][Nz][M], phi: [M]
    P = capon_power(xs.reshape(F * Nx, Nz, M), d, 32, 2)
    loss = -torch.log(P).mean()
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import CaponDesc, check, lib
from .capon import CaponPlan

_plans = {}


def _plan() -> CaponPlan:
    dev = torch.cuda.current_device()
    if dev not in _plans:
        _plans[dev] = CaponPlan()
    return _plans[dev]


def _desc(x, d, L, Yavg):
    Nx, Nz, M = x.shape
    ds = CaponDesc()
    ds.Nx, ds.Nz, ds.Mstride, ds.M, ds.L, ds.off, ds.Yavg = Nx, Nz, M, M, int(L), 0, int(Yavg)
    ds.d = float(d)
    return ds


class CaponPower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, d, L, Yavg):
        if not (x.is_cuda and x.dtype == torch.complex64):
            raise ValueError("x must be a complex64 CUDA tensor")
        x = x.contiguous()
        dval = float(d.detach().item()) if torch.is_tensor(d) else float(d)
        P = _plan().run(x, dval, L, Yavg, want=("P",))["P"]
        ctx.save_for_backward(x)
        ctx.meta = (dval, int(L), int(Yavg), torch.is_tensor(d))
        return P

    @staticmethod
    def backward(ctx, gP):
        (x,) = ctx.saved_tensors
        dval, L, Yavg, d_is_tensor = ctx.meta
        gP = gP.contiguous().to(torch.float32)
        gx = torch.zeros_like(x)
        gd = torch.empty_like(gP)
        ds = _desc(x, dval, L, Yavg)
        stream = torch.cuda.current_stream(x.device).cuda_stream
        check(lib().cos_capon_backward(_plan()._h, C.byref(ds), x.data_ptr(), gP.data_ptr(), gx.data_ptr(),
                                       gd.data_ptr(), stream))
        gdd = gd.sum().to(torch.float32) if (d_is_tensor and ctx.needs_input_grad[1]) else None
        return gx, gdd, None, None


def capon_power(x, d, L: int, Yavg: int):
    """Differentiable P [Nx*(Nz-2Yavg)] for x [Nx][Nz][M] (complex64, CUDA)."""
    return CaponPower.apply(x, d, L, Yavg)
