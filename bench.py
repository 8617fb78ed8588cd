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
    model = "unknown"
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": rate, "unit": "pixels/s", "cores": threads, "cpu_model": model, "kind": "port",
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
    ap.add_argument("--workload", default="frame",
                    choices=["frame", "aperture", "beamspace", "steering", "adjoint", "tdib", "lca", "solver", "shard"],
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
    P32 = out["P"].clone()          # the timed frame's outputs, checked below
    w32 = out["w"].clone()

    # dominant kernel alone (FP32 fused kernel, no fallback pass)
    kern_ms = timed(lambda: step("fp32", graph=False), max(5, args.steps // 2))
    achieved_tflops = FLOP_PER_PX * NPX / (kern_ms * 1e-3) / 1e12

    # accuracy: every pixel of the timed FP32 frame against the plain FP64 GPU kernel on the
    # same frame (the reference's own GPU-vs-FP64 comparison, testRTC.pyx:359-384); the
    # FP64 kernel itself is anchored to the FP64 oracle on a sample that includes every
    # scatterer's neighbourhood
    out64 = {k: torch.empty_like(v) for k, v in out.items()}
    fp64_ms = timed(lambda: plan.run(x, D, L, Y, precision="fp64", out=out64, stream=stream.cuda_stream), 2)
    P64 = out64["P"].double()
    w64 = out64["w"]
    Nzp = NZ - 2 * Y
    scat = synth.scatterer_positions(NX, NZ, seed=2)
    near = torch.zeros(NPX, dtype=torch.bool, device=dev)
    for zs, xs in scat:      # pixels whose centre row is within 8 rows of a scatterer, same beam
        lo, hi = max(zs - 8 - Y, 0), min(zs + 8 - Y + 1, Nzp)
        near[xs * Nzp + lo:xs * Nzp + hi] = True

    def errors(P, w=None):
        e = (P.double() - P64).abs() / P64
        r = {"P_rel_err_max": float(e.max()), "P_rel_err_p999": float(torch.quantile(e.float(), 0.999)),
             "P_rel_err_mean": float(e.mean()), "P_rel_err_max_near_scatterers": float(e[near].max()),
             "P_rel_err_p999_near_scatterers": float(torch.quantile(e[near].float(), 0.999))}
        if w is not None:
            we = torch.linalg.vector_norm(w - w64, dim=-1).double() / torch.linalg.vector_norm(w64, dim=-1).double()
            r.update({"w_rel_err_max": float(we.max()), "w_rel_err_p999": float(torch.quantile(we.float(), 0.999)),
                      "w_rel_err_max_near_scatterers": float(we[near].max())})
        return r

    acc = errors(P32, w32)
    acc["sample"] = (f"all {NPX} pixels vs the FP64 GPU kernel on the same frame; near scatterers: "
                     f"{int(near.sum())} pixels within 8 rows of the 10 scatterers (w: ||w - w64|| / ||w64||)")
    import oracle
    anchor_rows = [(0, NX, 900, 916)] + [(xs, xs + 1, max(zs - 8 - Y, 0), min(zs + 9 - Y, Nzp)) for zs, xs in scat]
    P64h = out64["P"].cpu().numpy().reshape(NX, Nzp)
    anchor = 0.0
    for x0, x1, j0, j1 in anchor_rows:
        ref = oracle.c_capon(x_np, D, L, Y, ix=(x0, x1), j=(j0, j1), want_w=False, nthreads=os.cpu_count() or 1)
        anchor = max(anchor, float((np.abs(P64h[x0:x1, j0:j1] - ref["P"]) / ref["P"]).max()))
    acc["fp64_kernel_vs_oracle_P_rel_err_max"] = anchor
    acc["fp64_kernel_anchor_sample"] = "256 beams x rows 900..915 and every scatterer neighbourhood, FP64 C oracle"
    res = {"fp64_plain_ms_per_frame": round(fp64_ms, 3), "speedup_vs_fp64_plain": round(fp64_ms / ms, 2)}
    # the same frame with one step of iterative refinement (residual from the snapshots)
    ref_ms = timed(lambda: step(refine=True, graph=False), max(3, args.steps // 4))
    res["refined"] = {"ms_per_frame": round(ref_ms, 3), "pixels_per_s": NPX / (ref_ms * 1e-3)}
    res["refined"].update(errors(out["P"], out["w"]))
    # the displacement recursion entirely on the VALU (the default forms its rank-two
    # corrections on the FP32 matrix cores)
    vr_ms = timed(lambda: step(cov_mfma="valu", graph=False), max(3, args.steps // 4))
    res["valu_recursion"] = {"ms_per_frame": round(vr_ms, 3), "pixels_per_s": NPX / (vr_ms * 1e-3),
                             "path": "covariance recursion with its rank-two corrections on the VALU"}
    res["valu_recursion"].update(errors(out["P"], out["w"]))
    # the covariance on the matrix cores (BF16 hi/lo split, FP32 accumulation) into a plan
    # buffer, then the FP32 solve from it
    mf_ms = timed(lambda: step(cov_mfma=True, graph=False), max(3, args.steps // 4))
    res["mfma_covariance"] = {"ms_per_frame": round(mf_ms, 3), "pixels_per_s": NPX / (mf_ms * 1e-3),
                              "path": "BF16 hi/lo MFMA covariance -> R in HBM -> FP32 Cholesky solve"}
    res["mfma_covariance"].update(errors(out["P"], out["w"]))

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
        "roofline": {"bound": "valu", "achieved": round(achieved_tflops, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
                     "traffic": measured_traffic(), "kernel": "capon_kernel<float,16,2,5,false,false,0,false,2>",
                     "kernel_ms": round(kern_ms, 4), "flop_per_pixel": FLOP_PER_PX},
        "accuracy": acc,
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
