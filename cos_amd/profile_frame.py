"""Runs the per-pixel near-field frame (256 beams x 2048 ranges x 64 elements, L = 32,
temporal window +/-2, loading 0.01 trace/L; seeded speckle and scatterers) a few times on
cuda:0, for a profiler to observe the kernels:

    python -m cos_amd.profile_frame [--cov default|valu|fused|rank|split] [--frames N] [--refine]
    python -m cos_amd.profile_frame --workload aperture [--frames N]

--cov selects the covariance: the library's default (the FP32 displacement recursion inside
the fused kernel, its rank-two corrections on the matrix cores where they apply), the
recursion entirely on the VALU (valu), per-snapshot Grams on the matrix cores inside the
fused kernel (fused), the rank-two corrections explicitly (rank), or the
separate matrix-core covariance into R followed by the solve (split).

--workload aperture runs the expanding-aperture frame instead (256 beams x 4096 ranges,
M(z) = 16..128, L(z) = M(z)/2, temporal window +/-2, 5% of the rows loaded at 1e-5 trace/L so
that they take the FP64 fallback; seeded), FP32 with the fallback pass.
"""
from __future__ import annotations

import argparse


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cov", default="default", choices=["default", "valu", "fused", "rank", "split"])
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--refine", action="store_true")
    ap.add_argument("--workload", default="frame", choices=["frame", "aperture"])
    args = ap.parse_args()
    import torch
    from cos_amd import CaponPlan, synth
    if args.workload == "aperture":
        x, Mz, Lz, off, dz = synth.config3_torch(256, 4096, seed=3, Mmax=128, device="cuda")
        rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
        plan = CaponPlan()
        for _ in range(args.frames):
            plan.run(x, 0.0, 0, 2, rows=rows, precision="fp32+fallback", refine=args.refine)
        torch.cuda.synchronize()
        print(f"{args.frames} expanding-aperture frames", flush=True)
        return
    x = torch.from_numpy(synth.config2(256, 2048, 64, seed=2)).cuda()
    plan = CaponPlan()
    mode = {"default": False, "valu": "valu", "fused": "fused", "rank": "rank", "split": True}[args.cov]
    for _ in range(args.frames):
        plan.run(x, 0.01, 32, 2, cov_mfma=mode, refine=args.refine)
    torch.cuda.synchronize()
    print(f"{args.frames} frames, covariance: {args.cov}", flush=True)


if __name__ == "__main__":
    main()
