"""Build libcos_capon.so in-tree with hipcc for gfx950.

    python -m cos_amd.build [--verbose]

Each source is compiled to an object next to nothing else (objects go to
cos_amd/_build/), then linked into cos_amd/libcos_capon.so.  All HIP code is
compiled with packed FP32 disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcos_capon.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = [
    "k_f32_g64c1_t5.hip", "k_f32_g64c1_t9.hip", "k_f32_g64c1_t0.hip",
    "k_f32_g16c2_t5.hip", "k_f32_g16c2_t5mg.hip", "k_f32_g16c2_t5mr.hip", "k_f32_g16c2_t5mw.hip",
    "k_f32_g16c2_t9.hip", "k_f32_g16c2_t0.hip",
    "k_f32_g16c1.hip", "k_f32_g16c3.hip", "k_f32_g16c3_t5mr.hip", "k_f32_ref_c1.hip", "k_f32_ref_c2.hip", "k_f32_ref_c3.hip", "k_f32_ref_g64.hip", "k_f32_ref5_c1.hip", "k_f32_ref5_c2.hip", "k_f32_ref5_c2mr.hip", "k_f32_ref5_c2mw.hip", "k_f32_ref5_c3.hip", "k_f32_ref5_c3mr.hip", "k_f32_ref5_g64.hip", "k_f64_c1.hip", "k_f64_c2.hip", "k_f64r_c1.hip", "k_f64r_c2.hip", "k_f64_g64.hip", "k_f64_g64_t5.hip", "k_bwd_c1.hip", "k_bwd_c2_m64.hip", "k_bwd_c2_m64mr.hip", "k_bwd_c2_m128.hip", "components.hip", "wide.hip", "cov_mfma.hip", "beamspace.hip", "tdib.hip", "lca.hip", "capon_dispatch.cpp", "capon_api.cpp",
]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        cmd += ["-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    if verbose and r.stderr:
        sys.stderr.write(r.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
