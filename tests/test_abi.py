"""The C ABI library loads, exports every declared symbol, and rejects bad arguments
the way the reference does -- all before any device call, so these run without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cos_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cos_capon.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cos_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def _err(code):
    return code, _lib.lib().cos_capon_last_error().decode()


@pytest.mark.parametrize("args,msg", [
    (dict(L=0), "L or M < 1"),
    (dict(Yavg=-1), "Yavg or Nb < 0"),
    (dict(Nb=9), "Nb can not be larger than L"),
    (dict(L=9, M=8), "L > M"),
    (dict(Yavg=2, Nz=4), "2*Yavg >= Nz"),
    (dict(L=73, M=80), "L > 72"),
])
def test_get_capon_argument_errors(args, msg):
    p = dict(d=0.1, L=4, Yavg=1, M=8, Nx=2, Nz=6, Nb=0)
    p.update(args)
    buf = np.zeros(8, np.complex64)
    code, text = _err(_lib.lib().cos_capon_get_capon(buf.ctypes.data, buf.ctypes.data, None, buf.ctypes.data,
                                                     p["d"], p["L"], p["Yavg"], p["M"], p["Nx"], p["Nz"], p["Nb"]))
    assert code != 0 and msg in text


def test_buildr_and_solver_argument_errors():
    L = _lib.lib()
    buf = np.zeros(8, np.complex64)
    code, text = _err(L.cos_buildr_get_r(buf.ctypes.data, buf.ctypes.data, 0.1, 4, 3, 8, 1, 6, 0, 0, 0))
    assert code != 0 and "Nz <= Yavg" in text
    code, text = _err(L.cos_solver_solve(buf.ctypes.data, buf.ctypes.data, buf.ctypes.data, 73, 1, 0, 0, 0))
    assert code != 0 and "N <= 72" in text


def test_plan_run_rejects_bad_descriptor():
    from cos_amd.capon import CaponDesc
    L = _lib.lib()
    h = C.c_void_p()
    assert L.cos_capon_plan_create(C.byref(h)) == 0
    ds = CaponDesc(Nx=1, Nz=8, Mstride=8, M=8, L=4, off=2, Yavg=1, d=0.1)
    code, text = _err(L.cos_capon_run(h, C.byref(ds), 16, None, None, None, 0, None, None, None))
    assert code != 0 and "off + M > Mstride" in text
    # the beamspace kernel's 32-bit row offsets bound the element stride
    ds = CaponDesc(Nx=1, Nz=8, Mstride=1 << 22, M=8, L=4, off=0, Yavg=1, d=0.1)
    code, text = _err(L.cos_capon_beamspace_run(h, C.byref(ds), 3, 16, None, None, None, None, 0, None))
    assert code != 0 and "Mstride < 4194304" in text
    # refinement is defined for the FP32 paths with the unsteered, unaveraged covariance
    for kw in (dict(precision=_lib.COS_PREC_FP64), dict(fb_avg=1)):
        ds = CaponDesc(Nx=1, Nz=8, Mstride=8, M=8, L=4, off=0, Yavg=1, d=0.1, refine=1, **kw)
        code, text = _err(L.cos_capon_run(h, C.byref(ds), 16, None, None, None, 0, None, None, None))
        assert code != 0 and "refinement applies to the FP32 paths" in text
    # the matrix-core covariance covers uniform L <= 32 with the default options
    for kw in (dict(L=65, M=70, Mstride=70), dict(refine=1), dict(p_mode=1)):
        args = dict(Nx=1, Nz=8, Mstride=64, M=64, L=16, off=0, Yavg=1, d=0.1, cov_mfma=1)
        args.update(kw)
        ds = CaponDesc(**args)
        code, text = _err(L.cos_capon_run(h, C.byref(ds), 16, None, None, None, 0, None, None, None))
        assert code != 0, kw
    assert L.cos_capon_plan_destroy(h) == 0


def test_tuning_table_rejects_shapes_too_small_for_L():
    from cos_amd.capon import CaponDesc
    L = _lib.lib()
    ds = CaponDesc(Nx=1, Nz=8, Mstride=64, M=64, L=32, off=0, Yavg=1, d=0.1)
    tab = (C.c_int * 65)()
    tab[20] = 1          # 16 lanes x 1 column holds L <= 16 only
    code, text = _err(L.cos_capon_tuning_set(C.byref(ds), tab))
    assert code != 0 and "does not hold its subarray length" in text
    tab[20] = 5          # no such shape
    code, text = _err(L.cos_capon_tuning_set(C.byref(ds), tab))
    assert code != 0
    tab[20] = 0
    tab[50] = 4          # 16 lanes x 3 columns holds L <= 48 only
    code, text = _err(L.cos_capon_tuning_set(C.byref(ds), tab))
    assert code != 0 and "does not hold its subarray length" in text


def test_shape_count_matches_header():
    from cos_amd.capon import N_SHAPES
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "cos_capon.h")).read()
    assert f"#define COS_CAPON_SHAPES {N_SHAPES}" in hdr


def test_getCaponCUDAPy_returns_error_code_like_reference():
    from cos_amd import getCaponCUDAPy
    x = np.zeros((2, 5, 4), np.complex64)
    z, R, w, e = getCaponCUDAPy(x, 0.1, 6, 1, np.array([0]), False, False)   # L > M
    assert e != 0
    assert z.shape == (2, 3) and w.shape == (2, 3, 6) and R.shape == (2, 3, 6, 6)


def test_every_export_has_argtypes():
    # ctypes passes a Python int without argtypes as a 32-bit int: pointers and stream
    # handles would be truncated
    L = _lib.lib()
    for name in _lib.EXPORTS:
        assert getattr(L, name).argtypes is not None, name
