"""The entry point that allocates nothing (cos_capon_run_noalloc, called through the C ABI)
against cos_capon_run on the same frames: identical P, z, w and flags, with the FP64
fallback pass running out of the library's fixed device scratch array.  Two calls in a
row, the second on a smaller frame (its list length must start again from zero), then a
frame past the array's pixel limit, which must be refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _noalloc(x, d, L, Y, **kw):
    import torch
    from cos_amd import CaponPlan
    from cos_amd._lib import check, lib
    ds, Lmax = CaponPlan._desc(x, d, L, Y, precision="fp32+fallback", **kw)
    Nx, Nz, _ = x.shape
    npx = Nx * (Nz - 2 * Y)
    o = {"P": torch.full((npx,), float("nan"), device="cuda"),
         "z": torch.full((npx,), float("nan"), dtype=torch.complex64, device="cuda"),
         "w": torch.zeros((npx, Lmax), dtype=torch.complex64, device="cuda"),
         "flags": torch.full((npx,), 255, dtype=torch.uint8, device="cuda")}
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().cos_capon_run_noalloc(C.byref(ds), x.data_ptr(), o["P"].data_ptr(), o["z"].data_ptr(),
                                      o["w"].data_ptr(), Lmax, o["flags"].data_ptr(), stream))
    torch.cuda.synchronize()
    return o


def _with_plan(x, d, L, Y, **kw):
    import torch
    from cos_amd import CaponPlan
    plan = CaponPlan()
    o = plan.run(x, d, L, Y, precision="fp32+fallback", want=("P", "z", "w", "flags"), **kw)
    torch.cuda.synchronize()
    return o, plan.fallback_count()


def _noalloc_no_flags(x, d, L, Y):
    """The same call with no flags output: the flags live in the scratch array too."""
    import torch
    from cos_amd import CaponPlan
    from cos_amd._lib import check, lib
    ds, Lmax = CaponPlan._desc(x, d, L, Y, precision="fp32+fallback")
    Nx, Nz, _ = x.shape
    npx = Nx * (Nz - 2 * Y)
    P = torch.full((npx,), float("nan"), device="cuda")
    check(lib().cos_capon_run_noalloc(C.byref(ds), x.data_ptr(), P.data_ptr(), None, None, 0, None,
                                      torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return P


def test_noalloc_matches_plan_entry_point():
    import torch
    from cos_amd import synth
    rng = np.random.default_rng(33)
    # call 1: two subarrays per window (M = 33, L = 32, Y = 0) and a loading of 1e-12, so
    # many FP32 factorisations break down and go to the FP64 pass through the scratch list
    M, L, Y, d = 33, 32, 0, 1e-12
    say("noalloc call 1", 8, 2048, M, L, Y, d)
    x1 = torch.from_numpy((rng.uniform(-1, 1, (8, 2048, M)) + 1j * rng.uniform(-1, 1, (8, 2048, M)))
                          .astype(np.complex64)).cuda()
    ref1, n1 = _with_plan(x1, d, L, Y, cond_thresh=1e30)
    o1 = _noalloc(x1, d, L, Y, cond_thresh=1e30)
    print(f"fallback pixels {n1} of {ref1['P'].numel()}", flush=True)
    assert n1 > 4096
    assert int((o1["flags"] != 0).sum()) == n1
    for k in ref1:
        assert torch.equal(o1[k], ref1[k]), k
    # call 2: a smaller frame, every pixel flagged by the condition bound (d = 1e-6)
    say("noalloc call 2", 4, 128, 64, 32, 2, 1e-6)
    x2 = torch.from_numpy(synth.config2(Nx=4, Nz=128, M=64, seed=2)).cuda()
    ref2, n2 = _with_plan(x2, 1e-6, 32, 2)
    o2 = _noalloc(x2, 1e-6, 32, 2)
    assert n2 == 4 * 124
    assert bool((o2["flags"] & 1).all())
    for k in ref2:
        assert torch.equal(o2[k], ref2[k]), k
    # call 3: no flags output (they go to the scratch array), the headline options
    say("noalloc call 3", 4, 128, 64, 32, 2, 0.01)
    P3 = _noalloc_no_flags(x2, 0.01, 32, 2)
    ref3, _ = _with_plan(x2, 0.01, 32, 2)
    assert torch.equal(P3, ref3["P"])
    P4 = _noalloc_no_flags(x2, 1e-6, 32, 2)
    assert torch.equal(P4, ref2["P"])


def test_noalloc_refuses_frames_past_its_scratch():
    import torch
    from cos_amd import CaponPlan
    from cos_amd._lib import lib
    say("noalloc size limit")
    x = torch.zeros((1, 8, 4), dtype=torch.complex64, device="cuda")
    ds, _ = CaponPlan._desc(x, 0.01, 2, 1, precision="fp32+fallback")
    ds.Nx, ds.Nz = 1025, 2048 + 2            # 1025 * 2048 output pixels > 2^21
    code = lib().cos_capon_run_noalloc(C.byref(ds), x.data_ptr(), None, None, None, 0, None, None)
    assert code != 0
    assert b"COS_CAPON_NOALLOC_MAX_PIXELS" in lib().cos_capon_last_error()
    # row tables are refused too
    rows = {k: torch.zeros(8, dtype=torch.int32, device="cuda") for k in ("L", "M", "off")}
    rows["d"] = torch.zeros(8, device="cuda")
    rows["L"] += 2
    rows["M"] += 4
    ds, _ = CaponPlan._desc(x, 0.01, 2, 1, precision="fp32+fallback", rows=rows)
    code = lib().cos_capon_run_noalloc(C.byref(ds), x.data_ptr(), None, None, None, 0, None, None)
    assert code != 0 and b"no row tables" in lib().cos_capon_last_error()
