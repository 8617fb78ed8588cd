"""cos_amd: MI355X-native Capon/MVDR adaptive beamformer (HIP kernels behind a C ABI).

The compute path is libcos_capon.so (built by ``python -m cos_amd.build``); this
package is its Python host side.  Importing it does not load the library; the
first call does, and raises if it is missing.
"""
from .capon import (BuildR, CaponPlan, Solver, getCaponCUDAPy, get_capon,  # noqa: F401
                    solveGaussJordanGPU)

__all__ = ["getCaponCUDAPy", "get_capon", "BuildR", "Solver", "solveGaussJordanGPU", "CaponPlan"]
