"""Python host side of the MI355X Capon beamformer.

Two layers, both over the C ABI of libcos_capon.so:

* Drop-in mirrors of the reference's Python/Cython boundary, numpy in and out:
  ``getCaponCUDAPy`` (framework/beamformer/capon/getCaponCUDA.pyx:10-81),
  ``BuildR.getR`` (RealTimeCapon/BuildR.pxd:8-25), ``Solver.solve``
  (RealTimeCapon/Solver.pxd:9-22) and ``solveGaussJordanGPU``
  (RealTimeCapon/Solver.pyx:27-57).  Same argument names, meanings, output
  layouts and the same integer error code ``e`` (0 = success).

* ``CaponPlan``: device tensors (torch, ROCm) in and out on the caller's stream,
  with the extensions the reference does not have: power image, per-row aperture
  tables, FP64 and FP64-fallback precision, steering vectors.

There is no CPU path: if the library is missing or the device call fails, these
functions raise (or return the nonzero code where the reference returns one).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import CaponDesc, check, lib

N_SHAPES = 4   # lane shapes of the fused FP32 kernel (COS_CAPON_SHAPES in include/cos_capon.h)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ------------------------------------------------------------------------------------
# Drop-in numpy layer
# ------------------------------------------------------------------------------------

def getCaponCUDAPy(Xd_in, d_in, L_in, Navg_in, V_in, doFBAvg=False, verbose=False):
    """Mirror of getCaponCUDAPy (CAP/getCaponCUDA.pyx:10-81).

    Xd_in [Nx][Ny][M] complex; returns [z [Nx][Ny-2Navg] c64, R [Nx][Ny-2Navg][L][L] c64
    (loaded, unnormalised, conj-stored), w [Nx][Ny-2Navg][L] c64, e (int error code)].
    Nb = V_in.shape[0] when V_in is 2-D (its contents are ignored, as in the reference);
    doFBAvg and verbose are accepted and ignored, as in the reference GPU path.
    """
    Nx, Ny, M = Xd_in.shape
    Navg, L = int(Navg_in), int(L_in)
    d = float(np.float32(d_in))
    Nb = V_in.shape[0] if getattr(V_in, "ndim", 0) == 2 else 0
    Npy = max(Ny - 2 * Navg, 0)
    z = np.zeros((Nx, Npy), dtype=np.complex64)
    w = np.zeros((Nx, Npy, L), dtype=np.complex64)
    R = np.zeros((Nx, Npy, L, L), dtype=np.complex64)
    Xd = np.ascontiguousarray(Xd_in, dtype=np.complex64)
    e = lib().cos_capon_get_capon(_ptr(z), _ptr(w), _ptr(R), _ptr(Xd), d, L, Navg, M, Nx, Ny, Nb)
    return [z, R, w, int(e)]


def get_capon(Xd, d, L, Yavg, Nb=0, want_R=False, want_w=True, want_z=True, want_P=True):
    """numpy convenience over cos_capon_get_capon_p: returns a dict of the requested
    outputs (P float32 power, z, w, R) and raises on error."""
    Nx, Nz, M = Xd.shape
    Npy = Nz - 2 * Yavg
    x = np.ascontiguousarray(Xd, dtype=np.complex64)
    out = {}
    if want_z:
        out["z"] = np.zeros((Nx, Npy), dtype=np.complex64)
    if want_w:
        out["w"] = np.zeros((Nx, Npy, L), dtype=np.complex64)
    if want_R:
        out["R"] = np.zeros((Nx, Npy, L, L), dtype=np.complex64)
    if want_P:
        out["P"] = np.zeros((Nx, Npy), dtype=np.float32)
    p = lambda k: _ptr(out[k]) if k in out else None
    check(lib().cos_capon_get_capon_p(p("z"), p("w"), p("R"), p("P"), _ptr(x), float(d), int(L),
                                      int(Yavg), M, Nx, Nz, int(Nb)))
    return out


class BuildR:
    """Mirror of the reference BuildR class (RealTimeCapon/BuildR.pxd:8-25)."""

    def getR(self, R, x, d, L, Yavg, M, Nx, Nz, Nb, R_on_gpu=False, x_on_gpu=False):
        """R, x: numpy complex64 arrays (host) or integer device addresses with the
        matching *_on_gpu flag.  Returns the integer error code."""
        Rp = int(R) if R_on_gpu else _ptr(R)
        if x_on_gpu:
            xp = int(x)
        else:
            xh = np.ascontiguousarray(x, dtype=np.complex64)
            xp = _ptr(xh)
        return int(lib().cos_buildr_get_r(Rp, xp, float(d), int(L), int(Yavg), int(M), int(Nx), int(Nz),
                                          int(Nb), int(bool(R_on_gpu)), int(bool(x_on_gpu))))


class Solver:
    """Mirror of the reference Solver class (RealTimeCapon/Solver.pxd:9-22): batched
    A x = b with column-major A."""

    def solve(self, x, A, b, N, batch, x_on_gpu=False, A_on_gpu=False, b_on_gpu=False):
        keep = []

        def host(arr):
            a = np.ascontiguousarray(arr, dtype=np.complex64)
            keep.append(a)
            return _ptr(a)

        xp = int(x) if x_on_gpu else _ptr(x)
        Ap = int(A) if A_on_gpu else host(A)
        bp = int(b) if b_on_gpu else host(b)
        return int(lib().cos_solver_solve(xp, Ap, bp, int(N), int(batch), int(bool(x_on_gpu)),
                                          int(bool(A_on_gpu)), int(bool(b_on_gpu))))


def solveGaussJordanGPU(A_in, b_in):
    """Mirror of RealTimeCapon/Solver.pyx:27-57: one system, A's C-order memory is
    handed over as is (so the column-major solver sees A^T).  Returns x [1][L]."""
    L = A_in.shape[0]
    A = np.ascontiguousarray(A_in, dtype=np.complex64)
    b = np.ascontiguousarray(b_in, dtype=np.complex64)
    x = np.zeros((1, L), dtype=np.complex64)
    Solver().solve(x, A, b, L, 1)
    return x


# ------------------------------------------------------------------------------------
# Device layer (torch tensors on the GPU)
# ------------------------------------------------------------------------------------

class CaponPlan:
    """Owns a cos_capon_plan (fallback bookkeeping buffers).  Not thread-safe: use one
    plan per thread."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().cos_capon_plan_create(C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            lib().cos_capon_plan_destroy(self._h)
            self._h = None

    def run(self, x, d, L, Yavg, *, M=None, off=0, precision="fp32", want=("P", "z", "w"),
            rows=None, steer_sin=None, cond_thresh=1e6, out=None, stream=None, graph=False,
            fb_avg=False, z_mode="amplitude", p_mode="loaded", refine=False, cov_mfma=False):
        """Run one frame.

        x: complex64 CUDA tensor [Nx][Nz][Mstride].  rows: optional dict of CUDA tensors
        L, M, off (int32) and d (float32), length Nz, indexed by the centre row.
        precision: "fp32", "fp64" or "fp32+fallback".  steer_sin: optional float32 CUDA
        tensor of sin(theta) values.  graph: capture the frame's launches in a hipGraph
        on the first call and replay it while the descriptor and all tensors are the same
        (pass the same ``out`` dict to keep the pointers).  fb_avg: forward-backward
        averaged covariance; z_mode: "amplitude" (w^H g) or "power" (power Capon);
        p_mode: "loaded" (1/(a^H R_L^-1 a)) or "unloaded" (|w^H R_hat w|).  refine: one step
        of iterative refinement on the FP32 paths (residual from the snapshots).  cov_mfma:
        True: covariance on the matrix cores (BF16 hi/lo split) into R, then the FP32 solve
        from it (L <= 64, row tables allowed, default options); "fused": per-snapshot Grams
        on the matrix cores inside the fused kernel where it applies (16 < L <= 32, windows
        of up to five rows, default options); "rank": the recursion's rank-two corrections
        summed per window on the FP32 matrix cores inside the fused kernel (same reach; what
        False does by default); "valu": the recursion entirely on the VALU.  Returns a dict
        of output tensors.
        """
        ds, o, ptr, wstride, stream = self._prepare(
            x, d, L, Yavg, M=M, off=off, precision=precision, want=want, rows=rows, steer_sin=steer_sin,
            cond_thresh=cond_thresh, out=out, stream=stream, graph=graph, fb_avg=fb_avg, z_mode=z_mode,
            p_mode=p_mode, refine=refine)
        ds.cov_mfma = {"fused": 2, "rank": 3, "valu": 4}.get(cov_mfma, cov_mfma) if cov_mfma not in (True, False) \
            else int(cov_mfma)
        check(lib().cos_capon_run(self._h, C.byref(ds), x.data_ptr(), ptr("P"), ptr("z"), ptr("w"),
                                  wstride, ptr("R"), ptr("flags"), stream))
        return o

    @staticmethod
    def _desc(x, d, L, Yavg, *, M=None, off=0, precision="fp32", rows=None, cond_thresh=1e6, graph=False,
              fb_avg=False, z_mode="amplitude", p_mode="loaded", refine=False):
        """The cos_capon_desc of a frame (see run())."""
        import torch

        if not (x.is_cuda and x.dtype == torch.complex64 and x.is_contiguous()):
            raise ValueError("x must be a contiguous complex64 CUDA tensor")
        Nx, Nz, Mstride = x.shape
        prec = {"fp32": _lib.COS_PREC_FP32, "fp64": _lib.COS_PREC_FP64,
                "fp32+fallback": _lib.COS_PREC_FP32_FALLBACK}[precision]
        ds = CaponDesc()
        ds.Nx, ds.Nz, ds.Mstride, ds.Yavg = Nx, Nz, Mstride, int(Yavg)
        ds.M = int(M if M is not None else Mstride)
        ds.L, ds.off, ds.d = int(L), int(off), float(d)
        ds.precision, ds.cond_thresh = prec, float(cond_thresh)
        ds.graph = int(bool(graph))
        ds.fb_avg = int(bool(fb_avg))
        ds.z_mode = {"amplitude": 0, "power": 1}[z_mode]
        ds.p_mode = {"loaded": 0, "unloaded": 1}[p_mode]
        ds.refine = int(bool(refine))
        Lmax = int(L)
        if rows is not None:
            for k in ("L", "M", "off", "d"):
                t = rows[k]
                if not (t.is_cuda and t.is_contiguous() and t.numel() == Nz):
                    raise ValueError(f"rows[{k}] must be a contiguous CUDA tensor of length Nz")
            ds.row_L, ds.row_M = rows["L"].data_ptr(), rows["M"].data_ptr()
            ds.row_off, ds.row_d = rows["off"].data_ptr(), rows["d"].data_ptr()
            Lmax = int(rows["L"].max().item())
            ds.Lmax, ds.Mmax = Lmax, int(rows["M"].max().item())
        return ds, Lmax

    def _prepare(self, x, d, L, Yavg, *, M, off, precision, want, rows, steer_sin, cond_thresh, out, stream,
                 graph, fb_avg, z_mode, p_mode, refine):
        """The frame descriptor and output tensors shared by run() and autotune()."""
        import torch

        ds, Lmax = CaponPlan._desc(x, d, L, Yavg, M=M, off=off, precision=precision, rows=rows,
                                   cond_thresh=cond_thresh, graph=graph, fb_avg=fb_avg, z_mode=z_mode,
                                   p_mode=p_mode, refine=refine)
        Nx, Nz, Mstride = x.shape
        S = 1
        if steer_sin is not None:
            ds.steer_sin, ds.nsteer = steer_sin.data_ptr(), int(steer_sin.numel())
            S = ds.nsteer
        npx = Nx * (Nz - 2 * Yavg)
        dev = x.device
        o = dict(out or {})
        if "P" in want and "P" not in o:
            o["P"] = torch.empty((npx * S,), dtype=torch.float32, device=dev)
        if "z" in want and "z" not in o:
            o["z"] = torch.empty((npx * S,), dtype=torch.complex64, device=dev)
        if "w" in want and "w" not in o:
            o["w"] = torch.zeros((npx * S, Lmax), dtype=torch.complex64, device=dev)
        if "R" in want and "R" not in o:
            o["R"] = torch.empty((npx, L * L), dtype=torch.complex64, device=dev)
        if "flags" in want and "flags" not in o:
            o["flags"] = torch.empty((npx,), dtype=torch.uint8, device=dev)
        ptr = lambda k: o[k].data_ptr() if k in o else None
        wstride = o["w"].shape[-1] if "w" in o else 0
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        return ds, o, ptr, wstride, stream

    def autotune(self, x, d, L, Yavg, *, M=None, off=0, rows=None, precision="fp32+fallback", fb_avg=False,
                 z_mode="amplitude", p_mode="loaded", refine=False, stream=None):
        """Time every lane shape of the fused FP32 kernel for each subarray length of this
        frame configuration and install the fastest for the device (later run() calls with
        the same Mmax, Yavg and options use it).  Arguments as run().  Returns
        {"shapes": {L: shape}, "ms": {L: {shape: ms per launch}}}; shape 1 = 16 lanes x 1
        column, 2 = 16 lanes x 2 columns, 3 = 64 lanes, 4 = 16 lanes x 3 columns (L <= 48)."""
        ds, o, ptr, wstride, stream = self._prepare(
            x, d, L, Yavg, M=M, off=off, precision=precision, want=("P", "z", "w"), rows=rows, steer_sin=None,
            cond_thresh=1e6, out=None, stream=stream, graph=False, fb_avg=fb_avg, z_mode=z_mode,
            p_mode=p_mode, refine=refine)
        shapes = (C.c_int * 65)()
        ms = (C.c_float * (65 * N_SHAPES))()
        check(lib().cos_capon_autotune(self._h, C.byref(ds), x.data_ptr(), ptr("P"), ptr("z"), ptr("w"),
                                       wstride, stream, shapes, ms))
        res = {"shapes": {}, "ms": {}}
        for l in range(1, 65):
            if shapes[l]:
                res["shapes"][l] = int(shapes[l])
                res["ms"][l] = {s + 1: float(ms[N_SHAPES * l + s]) for s in range(N_SHAPES)
                                if ms[N_SHAPES * l + s] >= 0}
        return res

    @staticmethod
    def set_tuning(x, d, L, Yavg, shapes, *, M=None, off=0, rows=None, precision="fp32+fallback", fb_avg=False,
                   z_mode="amplitude", p_mode="loaded", refine=False):
        """Install a lane-shape table {L: shape} (e.g. autotune()["shapes"], saved and
        reloaded) for the device and the frame configuration described as for run()."""
        ds, _ = CaponPlan._desc(x, d, L, Yavg, M=M, off=off, precision=precision, rows=rows, fb_avg=fb_avg,
                                z_mode=z_mode, p_mode=p_mode, refine=refine)
        tab = (C.c_int * 65)()
        for l, sh in shapes.items():
            tab[int(l)] = int(sh)
        check(lib().cos_capon_tuning_set(C.byref(ds), tab))

    def run_beamspace(self, x, d, L, Yavg, Nb, *, M=None, off=0, want=("P", "z"), out=None, stream=None):
        """Beamspace Capon with the sliding DFT over Nb Butler beams (loading in beamspace).
        x: complex64 CUDA tensor [Nx][Nz][Mstride].  want: any of P, z, w_bs, w (element
        space).  Returns a dict of output tensors."""
        import torch

        if not (x.is_cuda and x.dtype == torch.complex64 and x.is_contiguous()):
            raise ValueError("x must be a contiguous complex64 CUDA tensor")
        Nx, Nz, Mstride = x.shape
        ds = CaponDesc()
        ds.Nx, ds.Nz, ds.Mstride, ds.Yavg = Nx, Nz, Mstride, int(Yavg)
        ds.M = int(M if M is not None else Mstride)
        ds.L, ds.off, ds.d = int(L), int(off), float(d)
        npx = Nx * (Nz - 2 * Yavg)
        dev = x.device
        o = dict(out or {})
        shapes = {"P": ((npx,), torch.float32), "z": ((npx,), torch.complex64),
                  "w_bs": ((npx, Nb), torch.complex64), "w": ((npx, L), torch.complex64)}
        for k in want:
            if k not in o:
                o[k] = torch.empty(shapes[k][0], dtype=shapes[k][1], device=dev)
        ptr = lambda k: o[k].data_ptr() if k in o else None
        wstride = o["w"].shape[-1] if "w" in o else 0
        if stream is None:
            stream = torch.cuda.current_stream(dev).cuda_stream
        check(lib().cos_capon_beamspace_run(self._h, C.byref(ds), int(Nb), x.data_ptr(), ptr("P"), ptr("z"),
                                            ptr("w_bs"), ptr("w"), wstride, stream))
        return o

    def run_subspace(self, x, d, L, Yavg, V, *, M=None, off=0, fb_avg=False, want=("P", "z", "w"), out=None,
                     stream=None):
        """Capon in the subspace of the rows of V (complex64 CUDA tensor [Nb][L]), loading in the
        subspace (the reference CPU path's subspace option).  x: complex64 CUDA tensor
        [Nx][Nz][Mstride].  want: any of P, z, w (w is [npx][Nb]).  Returns a dict of tensors."""
        import torch

        if not (x.is_cuda and x.dtype == torch.complex64 and x.is_contiguous()):
            raise ValueError("x must be a contiguous complex64 CUDA tensor")
        if not (V.is_cuda and V.dtype == torch.complex64 and V.is_contiguous() and V.dim() == 2 and V.shape[1] == L):
            raise ValueError("V must be a contiguous complex64 CUDA tensor [Nb][L]")
        Nx, Nz, Mstride = x.shape
        Nb = V.shape[0]
        ds = CaponDesc()
        ds.Nx, ds.Nz, ds.Mstride, ds.Yavg = Nx, Nz, Mstride, int(Yavg)
        ds.M = int(M if M is not None else Mstride)
        ds.L, ds.off, ds.d = int(L), int(off), float(d)
        ds.fb_avg = int(bool(fb_avg))
        npx = Nx * (Nz - 2 * Yavg)
        o = dict(out or {})
        shapes = {"P": ((npx,), torch.float32), "z": ((npx,), torch.complex64), "w": ((npx, Nb), torch.complex64)}
        for k in want:
            if k not in o:
                o[k] = torch.empty(shapes[k][0], dtype=shapes[k][1], device=x.device)
        ptr = lambda k: o[k].data_ptr() if k in o else None
        if stream is None:
            stream = torch.cuda.current_stream(x.device).cuda_stream
        check(lib().cos_capon_subspace_run(self._h, C.byref(ds), V.data_ptr(), int(Nb), x.data_ptr(), ptr("P"),
                                           ptr("z"), ptr("w"), stream))
        return o

    def fallback_count(self) -> int:
        c = C.c_int()
        check(lib().cos_capon_last_fallback_count(self._h, C.byref(c)))
        return c.value
