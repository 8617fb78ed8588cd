"""ctypes binding of libcos_capon.so (the C ABI declared in include/cos_capon.h).

Loading fails loudly: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcos_capon.so")

# Every symbol declared in include/cos_capon.h.
EXPORTS = (
    "cos_capon_get_capon", "cos_capon_get_capon_p", "cos_buildr_get_r", "cos_solver_solve",
    "cos_capon_last_error", "cos_capon_plan_create", "cos_capon_plan_destroy", "cos_capon_run",
    "cos_capon_last_fallback_count", "cos_capon_beamspace_run", "cos_capon_backward",
    "cos_capon_subspace_run", "cos_tdib_run", "cos_lca_run", "cos_capon_autotune",
    "cos_capon_tuning_set", "cos_capon_run_noalloc", "cos_tdib_run_axis",
)

COS_PREC_FP32 = 0
COS_PREC_FP64 = 1
COS_PREC_FP32_FALLBACK = 2


class CaponDesc(C.Structure):
    """Mirror of cos_capon_desc."""
    _fields_ = [
        ("Nx", C.c_int), ("Nz", C.c_int), ("Mstride", C.c_int),
        ("M", C.c_int), ("L", C.c_int), ("off", C.c_int), ("Yavg", C.c_int),
        ("d", C.c_float),
        ("row_L", C.c_void_p), ("row_M", C.c_void_p), ("row_off", C.c_void_p), ("row_d", C.c_void_p),
        ("Lmax", C.c_int), ("Mmax", C.c_int),
        ("precision", C.c_int), ("cond_thresh", C.c_float),
        ("steer_sin", C.c_void_p), ("nsteer", C.c_int), ("graph", C.c_int),
        ("fb_avg", C.c_int), ("z_mode", C.c_int), ("p_mode", C.c_int), ("refine", C.c_int), ("cov_mfma", C.c_int),
    ]


_lib = None


class CosError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cos_capon error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m cos_amd.build`")
    # PyTorch-ROCm ships its own libamdhip64.so.7.  Loading torch first makes the
    # dynamic linker satisfy this library's libamdhip64.so.7 dependency with that same
    # runtime, so device pointers and streams from torch are valid here and only one
    # HIP runtime owns the GPU in the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, ip, fl = C.c_void_p, C.c_int, C.c_float
    L.cos_capon_get_capon.argtypes = [vp, vp, vp, vp, fl, ip, ip, ip, ip, ip, ip]
    L.cos_capon_get_capon_p.argtypes = [vp, vp, vp, vp, vp, fl, ip, ip, ip, ip, ip, ip]
    L.cos_buildr_get_r.argtypes = [vp, vp, fl, ip, ip, ip, ip, ip, ip, ip, ip]
    L.cos_solver_solve.argtypes = [vp, vp, vp, ip, ip, ip, ip, ip]
    L.cos_capon_last_error.restype = C.c_char_p
    L.cos_capon_last_error.argtypes = []
    L.cos_capon_plan_create.argtypes = [C.POINTER(vp)]
    L.cos_capon_plan_destroy.argtypes = [vp]
    L.cos_capon_run.argtypes = [vp, C.POINTER(CaponDesc), vp, vp, vp, vp, ip, vp, vp, vp]
    L.cos_capon_last_fallback_count.argtypes = [vp, C.POINTER(C.c_int)]
    L.cos_capon_run_noalloc.argtypes = [C.POINTER(CaponDesc), vp, vp, vp, vp, ip, vp, vp]
    L.cos_capon_beamspace_run.argtypes = [vp, C.POINTER(CaponDesc), ip, vp, vp, vp, vp, vp, ip, vp]
    L.cos_capon_backward.argtypes = [vp, C.POINTER(CaponDesc), vp, vp, vp, vp, vp]
    L.cos_capon_subspace_run.argtypes = [vp, C.POINTER(CaponDesc), vp, ip, vp, vp, vp, vp, vp]
    L.cos_capon_autotune.argtypes = [vp, C.POINTER(CaponDesc), vp, vp, vp, vp, ip, vp,
                                     C.POINTER(C.c_int), C.POINTER(C.c_float)]
    L.cos_capon_tuning_set.argtypes = [C.POINTER(CaponDesc), C.POINTER(C.c_int)]
    db = C.c_double
    L.cos_tdib_run.argtypes = [vp, ip, ip, db, db, vp, vp, db, db, vp, vp, ip, ip, fl, vp, vp, vp]
    L.cos_tdib_run_axis.argtypes = [vp, ip, ip, vp, vp, vp, db, db, vp, vp, ip, ip, fl, vp, vp, vp]
    L.cos_lca_run.argtypes = [vp, ip, ip, ip, vp, ip, ip, ip, vp, vp, vp]
    for name in EXPORTS:
        getattr(L, name).restype = getattr(L, name).restype or C.c_int
    L.cos_capon_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(code: int) -> None:
    if code != 0:
        msg = lib().cos_capon_last_error().decode(errors="replace")
        raise CosError(code, msg)
