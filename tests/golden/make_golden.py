"""Write the reference's own known-answer vectors as JSON fixtures.

Run once in a checkout that has the reference tree (default /root/reference);
the GPU box and the tests only read the committed JSON.  Nothing here executes
reference code: TestData.h comment blocks are parsed as text, and the Python
fixtures are the list literals of LA/tests/testMatrixOperations.py read with
ast.literal_eval.

Sources (paths relative to the reference root):
  apps/RealTimeCaponTest/TestData.h:7-25,67-85     inputs testRightside, testX16
  apps/RealTimeCaponTest/TestData.h:87-95          R normalised by 1/K, L=8, x=1..16 (4 dp)
  apps/RealTimeCaponTest/TestData.h:97-105         R unnormalised, L=8, x=1..16 (exact)
  apps/RealTimeCaponTest/TestData.h:107-117        Yavg=1 over [ones, x16, x16], loaded (x1e3, 4 dp)
  apps/RealTimeCaponTest/TestData.h:123-173,176-192 M=32 input [1..16,1..16], L=16 (exact)
  framework/linalg/tests/testMatrixOperations.py:16-67  Cholesky/solve/Butler vectors
"""
from __future__ import annotations

import ast
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _matrix_after(lines, anchor, n, skip=0, scale=1.0):
    i = next(k for k, s in enumerate(lines) if anchor in s)
    rows = []
    k = i + 1 + skip
    while len(rows) < n:
        s = lines[k].strip()
        k += 1
        if not s.startswith("//"):
            continue
        nums = re.findall(r"-?\d+(?:\.\d+)?", s[2:])
        if len(nums) == n:
            rows.append([float(v) * scale for v in nums])
    return rows


def _cucomplex_array(text, name):
    m = re.search(r"cuComplex\s+%s\s*\[[^\]]*\]\s*=\s*\{(.*?)\};" % name, text, re.S)
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    pairs = re.findall(r"make_cuComplex\(\s*([-0-9.eE+f]+)\s*,\s*([-0-9.eE+f]+)\s*\)", body)
    return [[float(a.rstrip("f")), float(b.rstrip("f"))] for a, b in pairs]


def _arith(node):
    """Evaluate a literal built from numbers, lists and + - * / (e.g. 3.0/2, [0.5]*4)."""
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, complex)):
        return node.value
    if isinstance(node, ast.List):
        return [_arith(e) for e in node.elts]
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _arith(node.operand)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _arith(node.left), _arith(node.right)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
    raise ValueError("not a plain arithmetic literal")


def _py_fixtures(path, names):
    tree = ast.parse(open(path).read())
    out = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and len(node.targets) == 1:
            t = node.targets[0]
            if isinstance(t, ast.Attribute) and t.attr in names and isinstance(node.value, ast.Call):
                try:
                    out[t.attr] = _arith(node.value.args[0])
                except (ValueError, IndexError):
                    continue
    return out


def _cplx(v):
    if isinstance(v, list):
        return [_cplx(e) for e in v]
    v = complex(v)
    return [v.real, v.imag]


def main(ref_root="/root/reference"):
    th = os.path.join(ref_root, "apps/RealTimeCaponTest/TestData.h")
    text = open(th).read()
    lines = text.splitlines()
    fx = {
        "source": "apps/RealTimeCaponTest/TestData.h; framework/linalg/tests/testMatrixOperations.py",
        "testRightside": _cucomplex_array(text, "testRightside"),
        "testX16": _cucomplex_array(text, "testX16"),
        "testX32_active": _cucomplex_array(text, "testX32"),
        "R8_normalised_4dp": _matrix_after(lines, "Resulting R after building (L=8)", 8),
        "R8_unnormalised": _matrix_after(lines, "Answer before normalization", 8),
        "R8_yavg1_loaded_4dp": _matrix_after(lines, "Answer after Yavg = 1", 8, scale=1e3),
        "R16_unnormalised": _matrix_after(lines, "Resulting R after building (L=16)", 16),
    }
    mo = os.path.join(ref_root, "framework/linalg/tests/testMatrixOperations.py")
    names = ["A", "R", "b1", "x1", "b2", "x2", "complexA", "complexR", "complexb",
             "complexy", "complexx", "but3", "but4", "AA", "complexAb"]
    py = _py_fixtures(mo, names)
    fx["linalg"] = {k: _cplx(v) for k, v in py.items()}
    out = os.path.join(HERE, "reference_fixtures.json")
    with open(out, "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", out, sorted(fx["linalg"]))


if __name__ == "__main__":
    main(*sys.argv[1:])
