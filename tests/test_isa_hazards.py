"""The fused kernels broadcast Cholesky multipliers inside FMAs (DPP source operands) and
insert the two wait states a DPP read needs after a VALU write of its source only where
the source was just written.  This test compiles the FP32 16-lane kernels to gfx950
assembly and checks, instruction by instruction, that no DPP instruction reads a VGPR
written fewer than two wait states earlier.  Runs on the CPU (hipcc cross-compiles)."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from cos_amd import build as B

HIPCC = B.HIPCC
SOURCES = ["k_f32_g16c2_t5.hip", "k_f32_g16c1.hip", "k_f32_g16c3.hip", "k_f32_ref_c2.hip", "k_f32_ref5_c2.hip", "k_f32_ref5_c3.hip", "k_bwd_c2_m64.hip"]

DPP = re.compile(r"^(v_\w+_dpp)\s+(v\d+),\s*-?\|?(v\d+)")
DST = re.compile(r"^v_\w+\s+(v\d+|v\[\d+:\d+\])")


def written(ins):
    m = DST.match(ins)
    if not (m and ins.startswith("v_")):
        return []
    d = m.group(1)
    if d.startswith("v["):
        lo, hi = d[2:-1].split(":")
        return [f"v{r}" for r in range(int(lo), int(hi) + 1)]
    return [d]


def hazards(lines):
    """DPP instructions whose src0 VGPR was written by a VALU instruction fewer than two
    wait states (instructions, or s_nop N = N + 1) before them."""
    ins = [l.strip() for l in lines]
    ins = [l for l in ins if l and re.match(r"^[a-z]", l) and not l.endswith(":")]
    found, checked = [], 0
    for n, l in enumerate(ins):
        m = DPP.match(l)
        if not m:
            continue
        checked += 1
        src, ws = m.group(3), 0
        for p in reversed(ins[max(0, n - 2):n]):
            if ws >= 2:
                break
            if p.startswith("s_nop"):
                ws += int(p.split()[1]) + 1
                continue
            if src in written(p):
                found.append((p, l))
                break
            ws += 1
    return found, checked


def test_checker_flags_a_planted_hazard():
    # a write right before the DPP read, one with s_nop 1 between, one with one instruction
    lines = ["v_mul_f32_e32 v5, v1, v2", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:3",
             "v_mul_f32_e32 v9, v1, v2", "s_nop 1", "v_fmac_f32_dpp v7, -v9, v8 row_newbcast:3",
             "v_mul_f32_e32 v10, v1, v2", "v_add_f32_e32 v11, v1, v2", "v_fmac_f32_dpp v7, -v10, v8 row_newbcast:3"]
    found, checked = hazards(lines)
    assert checked == 3 and [f[0] for f in found] == ["v_mul_f32_e32 v5, v1, v2", "v_mul_f32_e32 v10, v1, v2"]


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fused_kernels_have_no_dpp_read_hazards():
    tmp = tempfile.mkdtemp(prefix="cos_isa_")
    try:
        procs = []
        for src in SOURCES:
            out = os.path.join(tmp, src + ".s")
            cmd = [HIPCC] + [a for a in B.COMMON if a != "-fPIC"] + ["-x", "hip", "-S", "--cuda-device-only",
                                                                     os.path.join(B.CSRC, src), "-o", out]
            procs.append((out, subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)))
        total = 0
        for out, p in procs:
            _, err = p.communicate(timeout=600)
            assert p.returncode == 0, err.decode()[-2000:]
            found, checked = hazards(open(out).readlines())
            assert not found, f"{os.path.basename(out)}: {found[:3]}"
            total += checked
        assert total > 10000   # the Cholesky's broadcast FMAs are there to check
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
