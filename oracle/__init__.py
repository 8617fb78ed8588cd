"""The oracle: CPU FP64 restatements of the reference's Capon path.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product package.

``capon_oracle`` is the numpy restatement; ``c_capon`` runs the C restatement
(oracle/capon_ref.c, built into oracle/liboracle_capon.so by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle_capon.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return _LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.oracle_capon_f64.argtypes = [C.c_void_p] + [C.c_int] * 5 + [C.c_double] + [C.c_int] * 4 + \
                                      [C.c_void_p] * 3 + [C.c_int]
        L.oracle_capon_f64.restype = C.c_int
        L.oracle_max_threads.restype = C.c_int
        _lib = L
    return _lib


def max_threads() -> int:
    return _load().oracle_max_threads()


def c_capon(x: np.ndarray, d: float, L: int, Yavg: int, ix=(0, None), j=(0, None), want_w=True,
            nthreads: int = 0):
    """C FP64 Capon over pixels ix in [ix0, ix1), j in [j0, j1) (j indexes output rows).
    Returns dict P [nix][nj], z, w [nix][nj][L]."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    Nx, Nz, M = x.shape
    ix0, ix1 = ix[0], Nx if ix[1] is None else ix[1]
    j0, j1 = j[0], Nz - 2 * Yavg if j[1] is None else j[1]
    nix, nj = ix1 - ix0, j1 - j0
    P = np.zeros((nix, nj))
    z = np.zeros((nix, nj), dtype=np.complex128)
    w = np.zeros((nix, nj, L), dtype=np.complex128) if want_w else None
    r = _load().oracle_capon_f64(x.ctypes.data, Nx, Nz, M, L, Yavg, float(d), ix0, ix1, j0, j1,
                                 P.ctypes.data, z.ctypes.data, w.ctypes.data if want_w else None,
                                 int(nthreads))
    if r != 0:
        raise ValueError("oracle_capon_f64: bad arguments")
    return dict(P=P, z=z, w=w)
