#!/usr/bin/env python3
"""Headline benchmark: Capon pixels/s at M=64, L=32 on one MI355X.

Workload: the per-pixel near-field frame, cube [Nx=256][Nz=2048][M=64] complex64
(speckle + 10 scatterers, seeded synthetic data), L=32 (K=33 subarrays), temporal
window +/-2, loading 0.01*trace/L.  One step = one frame through the fused kernel
(power, amplitude and weights for all 256*2044 pixels) plus the FP64 fallback pass,
with the cube already resident in HBM.

Prints one JSON line (see the keys below).  --gpus > 1 is refused: this build runs on
one GPU (pixels are independent; the multi-GPU layout is documented, not launched).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX, NZ, M, L, Y, D = 256, 2048, 64, 32, 2, 0.01
K = M - L + 1
NPX = NX * (NZ - 2 * Y)
# Algorithmic work per output pixel (Gram once per snapshot reused along range,
# window add/sub, Cholesky, two triangular solves, outputs) and bytes per pixel
# (cube read once, P, w, z written).
FLOP_PER_PX = 8 * L * L * K + 4 * L * L + (4 * L ** 3) // 3 + 8 * L * L + 32 * L   # 327,339
BYTES_PER_PX = 8 * M + 4 + 8 * L + 8                                                # 780
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def measured_traffic():
    """HBM bytes per dispatch of the fused kernel from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "capon_fused_f32_cfg2_traffic.txt")
    try:
        kv = dict(l.strip().split("=", 1) for l in open(path) if "=" in l and not l.startswith("#"))
        return int(kv["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def end_to_end(x, calls=3):
    """The drop-in call with host buffers (cos_capon_get_capon_p: the library copies the
    cube in and z, w, P out over PCIe, R not requested), wall time per call after one
    warm-up call.  Reported beside the resident-cube figure, never as `value`."""
    import ctypes as C
    from cos_amd._lib import lib
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    xc = np.ascontiguousarray(x, dtype=np.complex64)
    z = np.empty((NX, NZ - 2 * Y), np.complex64)
    w = np.empty((NX, NZ - 2 * Y, L), np.complex64)
    P = np.empty((NX, NZ - 2 * Y), np.float32)
    call = lambda: lib().cos_capon_get_capon_p(ptr(z), ptr(w), None, ptr(P), ptr(xc), D, L, Y, M, NX, NZ, 0)
    if call() != 0:
        return None
    t0 = time.perf_counter()
    for _ in range(calls):
        call()
    ms = (time.perf_counter() - t0) / calls * 1e3
    host_bytes = xc.nbytes + z.nbytes + w.nbytes + P.nbytes
    return {"ms_per_frame": round(ms, 3), "pixels_per_s": NPX / (ms * 1e-3),
            "host_bytes": host_bytes, "calls": calls,
            "path": "cos_capon_get_capon_p, pageable host buffers: cube in; z, w, P out; R not requested"}


def cpu_baseline(x, budget_s=12.0):
    """The oracle's C FP64 restatement of the reference CPU path, all host cores, on a
    bounded sample of whole beams; rate scaled to pixels/s."""
    import oracle
    threads = os.cpu_count() or 1
    oracle.c_capon(x[:1, :64], D, L, Y, nthreads=threads)           # warm up / load
    # all host threads: whole frames, repeated until about budget_s of work
    t0 = time.perf_counter()
    frames = 0
    while True:
        oracle.c_capon(x, D, L, Y, nthreads=threads, want_w=True)
        frames += 1
        t = time.perf_counter() - t0
        if t > budget_s or t * (frames + 1) / frames > 2.5 * budget_s:
            break
    rate = frames * NPX / t
    # one thread (the reference CPU path is single-threaded), on two beams
    t0 = time.perf_counter()
    oracle.c_capon(x, D, L, Y, ix=(0, 2), nthreads=1, want_w=True)
    t1 = time.perf_counter() - t0
    return {"value": rate, "unit": "pixels/s", "cores": threads, "kind": "port",
            "sample": f"{frames} whole frame(s) (FP64, OpenMP over pixels) in {t:.1f} s",
            "single_thread_value": 2 * (NZ - 2 * Y) / t1,
            "single_thread_sample": f"2 beams x {NZ - 2 * Y} rows on 1 thread, {t1:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a frame takes ~1.7 ms: 50 untimed frames bring the clocks to steady state before the
    # 100 timed ones (0.25 s of GPU time in all)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparator")
    ap.add_argument("--workload", default="frame",
                    choices=["frame", "aperture", "beamspace", "steering", "adjoint", "tdib", "lca", "shard"],
                    help="frame: the headline 256x2048x64 L=32 frame (default); the others time the "
                         "expanding-aperture, sonar beamspace, steered and differentiable workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); pixels shard by beam "
              "with no exchange, so N GPUs would each run this frame on Nx/N beams.", file=sys.stderr)
        sys.exit(2)
    if args.workload != "frame":
        import bench_workloads
        print(json.dumps(getattr(bench_workloads, args.workload)(args)), flush=True)
        return

    import torch
    from cos_amd import CaponPlan, synth

    t0 = time.perf_counter()
    x_np = synth.config2(NX, NZ, M, seed=2)
    gen_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    x = torch.from_numpy(x_np).to(dev)
    plan = CaponPlan()
    out = {
        "P": torch.empty(NPX, dtype=torch.float32, device=dev),
        "z": torch.empty(NPX, dtype=torch.complex64, device=dev),
        "w": torch.empty((NPX, L), dtype=torch.complex64, device=dev),
    }
    stream = torch.cuda.current_stream(dev)

    def step(prec="fp32+fallback", graph=True, refine=False, cov_mfma=False):
        # one frame; the launches are captured once in a hipGraph and replayed
        plan.run(x, D, L, Y, precision=prec, out=out, stream=stream.cuda_stream, graph=graph, refine=refine,
                 cov_mfma=cov_mfma)

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n          # ms

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ms = timed(step, args.steps)
    wall = time.perf_counter() - wall0
    nfb = plan.fallback_count()

    # dominant kernel alone (FP32 fused kernel, no fallback pass)
    kern_ms = timed(lambda: step("fp32", graph=False), max(5, args.steps // 2))
    achieved_tflops = FLOP_PER_PX * NPX / (kern_ms * 1e-3) / 1e12

    # accuracy on a sample vs the FP64 oracle, and the plain FP64 GPU kernel
    import oracle
    P = out["P"].cpu().numpy().reshape(NX, NZ - 2 * Y)
    Wt = out["w"].cpu().numpy().reshape(NX, NZ - 2 * Y, L)[:, 900:916]
    ref = oracle.c_capon(x_np, D, L, Y, ix=(0, NX, ), j=(900, 916), want_w=True,
                         nthreads=os.cpu_count() or 1)
    err = np.abs(P[:, 900:916] - ref["P"]) / ref["P"]
    werr = np.linalg.norm(Wt - ref["w"], axis=-1) / np.linalg.norm(ref["w"], axis=-1)
    res = {}
    if not args.no_fp64:
        fp64_ms = timed(lambda: step("fp64", graph=False), 2)
        res["fp64_plain_ms_per_frame"] = round(fp64_ms, 3)
        res["speedup_vs_fp64_plain"] = round(fp64_ms / ms, 2)
    # the same frame with one step of iterative refinement (residual from the snapshots)
    ref_ms = timed(lambda: step(refine=True, graph=False), max(3, args.steps // 4))
    Pr = out["P"].cpu().numpy().reshape(NX, NZ - 2 * Y)
    err_r = np.abs(Pr[:, 900:916] - ref["P"]) / ref["P"]
    res["refined"] = {"ms_per_frame": round(ref_ms, 3), "pixels_per_s": NPX / (ref_ms * 1e-3),
                      "P_rel_err_max": float(err_r.max()), "P_rel_err_mean": float(err_r.mean()),
                      "unrefined_P_rel_err_mean": float(err.mean())}
    # the covariance on the matrix cores (BF16 hi/lo split, FP32 accumulation) into a plan
    # buffer, then the FP32 solve from it
    mf_ms = timed(lambda: step(cov_mfma=True, graph=False), max(3, args.steps // 4))
    Pm = out["P"].cpu().numpy().reshape(NX, NZ - 2 * Y)
    err_m = np.abs(Pm[:, 900:916] - ref["P"]) / ref["P"]
    res["mfma_covariance"] = {"ms_per_frame": round(mf_ms, 3), "pixels_per_s": NPX / (mf_ms * 1e-3),
                              "P_rel_err_max": float(err_m.max()), "P_rel_err_mean": float(err_m.mean()),
                              "path": "BF16 hi/lo MFMA covariance -> R in HBM -> FP32 Cholesky solve"}

    line = {
        "metric": "Capon pixels/sec at M=64,L=32 (<=1e-4 rel err vs FP64); % HBM/MFMA roofline",
        "value": NPX / (ms * 1e-3),
        "unit": "pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded speckle + 10 point scatterers at 30 dB)",
        "config": {"workload": "per-pixel near-field Capon frame", "Nx": NX, "Nz": NZ, "M": M, "L": L,
                   "Yavg": Y, "d": D, "pixels": NPX, "precision": "fp32 + fp64 fallback",
                   "fallback_pixels": nfb},
        "roofline": {"bound": "mfma", "achieved": round(achieved_tflops, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
                     "traffic": measured_traffic(), "kernel": "capon_kernel<float,16,2,5,false,false,0>",
                     "kernel_ms": round(kern_ms, 4), "flop_per_pixel": FLOP_PER_PX},
        "accuracy": {"P_rel_err_max": float(err.max()), "P_rel_err_p999": float(np.quantile(err, 0.999)),
                     "w_rel_err_max": float(werr.max()), "w_rel_err_mean": float(werr.mean()),
                     "sample": "256 beams x rows 900..915 vs FP64 oracle (w: ||w - w64|| / ||w64||)"},
        "wall_s_timed": round(wall, 3),
        "input_gen_s": round(gen_s, 2),
    }
    line.update(res)
    line["end_to_end"] = end_to_end(x_np)
    if not args.no_cpu:
        cb = cpu_baseline(x_np)
        line["cpu_baseline"] = cb
        line["speedup_vs_cpu"] = round(line["value"] / cb["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
