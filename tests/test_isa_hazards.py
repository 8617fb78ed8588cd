"""The fused kernels broadcast Cholesky multipliers inside FMAs (DPP source operands) and
insert the two wait states a DPP read needs after a VALU write of its source only where
the source was just written.  This test compiles the FP32 16-lane kernels to gfx950
assembly and checks that no DPP instruction reads a VGPR written fewer than two wait
states earlier on any path into it: the walk back crosses block boundaries into every
predecessor (loop back edges, branches, fall-through).  Runs on the CPU (hipcc
cross-compiles)."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from cos_amd import build as B

HIPCC = B.HIPCC
# every FP32 16-lane kernel (the ones with DPP-broadcast FMAs)
SOURCES = ["k_f32_g16c2_t5mr.hip", "k_f32_g16c2_t5mw.hip", "k_f32_g16c2_t5.hip", "k_f32_g16c2_t9.hip", "k_f32_g16c2_t0.hip",
           "k_f32_g16c2_t5mg.hip", "k_f32_g16c1.hip", "k_f32_g16c3.hip", "k_f32_g16c3_t5mr.hip", "k_f32_ref_c1.hip", "k_f32_ref_c2.hip",
           "k_f32_ref_c3.hip", "k_f32_ref5_c1.hip", "k_f32_ref5_c2.hip", "k_f32_ref5_c2mr.hip", "k_f32_ref5_c2mw.hip",
           "k_f32_ref5_c3.hip", "k_f32_ref5_c3mr.hip", "k_bwd_c1.hip", "k_bwd_c2_m64.hip", "k_bwd_c2_m64mr.hip",
           "k_bwd_c2_m128.hip"]

DPP = re.compile(r"^(v_\w+_dpp)\s+(v\d+),\s*-?\|?(v\d+)")
DST = re.compile(r"^v_\w+\s+(v\d+|v\[\d+:\d+\])")
SWAP = re.compile(r"^v_permlane(16|32)_swap\w*\s+(v\d+),\s*(v\d+)")


def written(ins):
    s = SWAP.match(ins)
    if s:   # the lane swaps write both their operands
        return [s.group(2), s.group(3)]
    m = DST.match(ins)
    if not (m and ins.startswith("v_")):
        return []
    d = m.group(1)
    if d.startswith("v["):
        lo, hi = d[2:-1].split(":")
        return [f"v{r}" for r in range(int(lo), int(hi) + 1)]
    return [d]


BRANCH = re.compile(r"^s_(c?branch\w*)\s+(\S+)")


def blocks(lines):
    """Basic blocks of an assembly listing: (instructions, successor labels, falls through).
    A block starts at a label or after a branch; s_branch and s_endpgm end without fall-through."""
    out, cur, label = [], [], None
    names = {}

    def close(succ, fall):
        nonlocal cur, label
        out.append({"label": label, "ins": cur, "succ": succ, "fall": fall})
        if label:
            names[label] = len(out) - 1
        cur, label = [], None

    for raw in lines:
        l = raw.split(";")[0].strip()
        if not l:
            continue
        if l.endswith(":") and not l.startswith("."):
            continue                      # function symbol
        if l.endswith(":"):
            if cur or label:
                close([], True)
            label = l[:-1]
            continue
        if not re.match(r"^[a-z]", l):
            continue                      # directives
        cur.append(l)
        m = BRANCH.match(l)
        if m:
            close([m.group(2)], not m.group(1).startswith("branch"))
        elif l.startswith("s_endpgm"):
            close([], False)
    if cur or label:
        close([], False)
    preds = [[] for _ in out]
    for i, b in enumerate(out):
        for t in b["succ"]:
            if t in names:
                preds[names[t]].append(i)
        if b["fall"] and i + 1 < len(out):
            preds[i + 1].append(i)
    return out, preds


def hazards(lines):
    """DPP instructions whose src0 VGPR was written by a VALU instruction fewer than two
    wait states (instructions, or s_nop N = N + 1) before them, on any path into them: the
    walk back continues through every predecessor block (a loop's back edge, a branch)."""
    bl, preds = blocks(lines)
    found, checked = [], 0

    def walk(b, end, src, ws, seen):
        ins = bl[b]["ins"]
        for p in reversed(ins[:end]):
            if ws >= 2:
                return None
            if p.startswith("s_nop"):
                ws += int(p.split()[1]) + 1
                continue
            if src in written(p):
                return p
            ws += 1
        if ws >= 2:
            return None
        for q in preds[b]:
            if (q, ws) in seen:
                continue
            hit = walk(q, len(bl[q]["ins"]), src, ws, seen | {(q, ws)})
            if hit:
                return hit
        return None

    for b, blk in enumerate(bl):
        for n, l in enumerate(blk["ins"]):
            m = DPP.match(l)
            if not m:
                continue
            checked += 1
            hit = walk(b, n, m.group(3), 0, frozenset())
            if hit:
                found.append((hit, l))
    return found, checked


def test_checker_flags_a_planted_hazard():
    # a write right before the DPP read, one with s_nop 1 between, one with one instruction
    lines = ["v_mul_f32_e32 v5, v1, v2", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:3",
             "v_mul_f32_e32 v9, v1, v2", "s_nop 1", "v_fmac_f32_dpp v7, -v9, v8 row_newbcast:3",
             "v_mul_f32_e32 v10, v1, v2", "v_add_f32_e32 v11, v1, v2", "v_fmac_f32_dpp v7, -v10, v8 row_newbcast:3"]
    found, checked = hazards(lines)
    assert checked == 3 and [f[0] for f in found] == ["v_mul_f32_e32 v5, v1, v2", "v_mul_f32_e32 v10, v1, v2"]


def test_checker_counts_both_operands_of_a_lane_swap():
    # v_permlane16_swap writes its second operand too; a DPP read of it right after waits
    lines = ["v_permlane16_swap_b32_e32 v4, v5", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:3",
             "v_permlane16_swap_b32_e32 v9, v10", "s_nop 1", "v_fmac_f32_dpp v7, -v10, v8 row_newbcast:3"]
    found, checked = hazards(lines)
    assert checked == 2 and [f[0] for f in found] == ["v_permlane16_swap_b32_e32 v4, v5"]


def test_checker_follows_loop_back_edges_and_branches():
    # the loop head reads v5 by DPP; along the back edge only the branch (one wait state)
    # separates it from the write
    loop = [".LBB0_1:", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:3", "v_add_f32_e32 v1, v1, v2",
            "v_mul_f32_e32 v5, v1, v2", "s_cbranch_scc1 .LBB0_1", "s_endpgm"]
    found, checked = hazards(loop)
    assert checked == 1 and [f[0] for f in found] == ["v_mul_f32_e32 v5, v1, v2"]
    # one more instruction between the write and the branch clears it
    ok = loop[:4] + ["s_add_u32 s0, s0, 1"] + loop[4:]
    assert hazards(ok)[0] == []
    # a forward branch into a block whose first instruction is the DPP read, from a block
    # that wrote the source last; the fall-through path is clean
    fwd = ["v_mul_f32_e32 v5, v1, v2", "s_cbranch_execz .LBB0_2", "s_nop 4", "v_mov_b32_e32 v9, v1",
           ".LBB0_2:", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:1", "s_endpgm"]
    found, _ = hazards(fwd)
    assert [f[0] for f in found] == ["v_mul_f32_e32 v5, v1, v2"]
    # an unconditional branch does not fall through into the next block
    jmp = ["v_mul_f32_e32 v5, v1, v2", "s_branch .LBB0_3", ".LBB0_2:", "v_fmac_f32_dpp v7, -v5, v8 row_newbcast:1",
           "s_endpgm", ".LBB0_3:", "s_endpgm"]
    assert hazards(jmp)[0] == []


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
