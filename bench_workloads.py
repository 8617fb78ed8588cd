"""Secondary workloads for bench.py (``--workload``): one JSON line each.

aperture   expanding aperture, 256 x 4096, M(z) 16..128, L(z) = M(z)/2, 5% of rows at
           d = 1e-5 (FP32 + FP64 fallback, against an all-FP64 run of the same frame)
beamspace  sonar beamspace, 512 look directions x 8192 ranges x 96 elements, L = 48,
           Nb = 5, Yavg = 4, power image (sliding-DFT kernel)
steering   far-field ULA, 1024 range bins x 32 elements, L = 16, 128 steering angles
adjoint    differentiable frames: 8 x (128 x 512 x 64) with per-element phase errors,
           loss = -mean log P, forward + adjoint
tdib       the delay stage producing a 256 x 2048 x 64 cube from 64 x 32768 samples
lca        window selection over 8 windows on a 256 x 2048 image
solver     Solver::solve on the headline frame's systems: 523,264 systems of N = 32
           (and 65,536 of N = 72), all operands resident on the device
shard      one GPU's share of the headline frame split by beam over 8 GPUs (32 beams)

Inputs are generated on the GPU with the counter-based generator (cos_amd.synth), so
they are resident in HBM before timing starts.  The oracle is used only for the accuracy
samples.
"""
from __future__ import annotations

import time

import numpy as np

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def _timed(torch, fn, n, stream=None):
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(n):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / n


def _host():
    """Threads the CPU legs use and the host's CPU model."""
    import os
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def _cpu_line(value, unit, sample, kind="port", extra=None):
    cores, model = _host()
    d = {"value": value, "unit": unit, "cores": cores, "cpu_model": model, "kind": kind, "sample": sample}
    if extra:
        d.update(extra)
    return d


def _flops(L, K):
    return 8 * L * L * K + 4 * L * L + (4 * L ** 3) // 3 + 8 * L * L + 32 * L


def _base(args, metric, value, unit, ms, cfg, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded)", "config": cfg}


def aperture(args):
    import torch
    from cos_amd import CaponPlan, synth
    from oracle import capon_oracle as O
    Nx, Nz, Mmax, Y = 256, 4096, 128, 2
    x, Mz, Lz, off, dz = synth.config3_torch(Nx, Nz, seed=3, Mmax=Mmax, device="cuda")
    rows = {k: torch.from_numpy(v).cuda() for k, v in dict(L=Lz, M=Mz, off=off, d=dz).items()}
    plan = CaponPlan()
    npx = Nx * (Nz - 2 * Y)
    out = {"P": torch.empty(npx, device="cuda"), "z": torch.empty(npx, dtype=torch.complex64, device="cuda"),
           "w": torch.empty((npx, 64), dtype=torch.complex64, device="cuda")}
    step = lambda refine=False, mf=False: plan.run(x, 0.0, 0, Y, rows=rows, precision="fp32+fallback", out=out,
                                                   refine=refine, cov_mfma=mf)
    for _ in range(args.warmup):
        step()
    ms_default = _timed(torch, step, args.steps)
    # the lane-shape autotuner: every shape that holds each L, timed on up to 128 rows of it
    t0 = time.perf_counter()
    tuned = plan.autotune(x, 0.0, 0, Y, rows=rows)
    tune_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    nfb = plan.fallback_count()
    P32 = out["P"].clone()
    o64 = {"P": torch.empty(npx, device="cuda")}
    ms64 = _timed(torch, lambda: plan.run(x, 0.0, 0, Y, rows=rows, precision="fp64", out=o64, want=("P",)), 1)
    e = (P32.double() - o64["P"].double()).abs() / o64["P"].double()
    c = np.arange(Nz - 2 * Y) + Y
    fpx = float(np.mean([_flops(int(Lz[cc]), int(Mz[cc] - Lz[cc] + 1)) for cc in c]))
    # oracle sample: 4 beams x 16 rows spread over the depth (FP64 numpy)
    rng = np.random.default_rng(3)
    beams = np.sort(rng.choice(Nx, 4, replace=False))
    ib, jj = np.repeat(np.arange(4), 16), rng.integers(0, Nz - 2 * Y, 64)
    xs = x[torch.from_numpy(beams).cuda()].cpu().numpy().astype(np.complex128)
    ref = O.capon_rows(xs, dz.astype(np.float64), Lz, Mz, off, Y, pixels=(ib, jj))
    Ps = P32.cpu().numpy().reshape(Nx, Nz - 2 * Y)[beams[ib], jj]
    eo = np.abs(Ps - ref["P"]) / ref["P"]
    Ws = out["w"].cpu().numpy().reshape(Nx, Nz - 2 * Y, 64)
    ew = [np.linalg.norm(Ws[beams[b], j, :len(w)] - w) / np.linalg.norm(w) for b, j, w in zip(ib, jj, ref["w"])]
    # one step of iterative refinement on the FP32 runs (FP64 fallback as before)
    ms_ref = _timed(torch, lambda: step(True), max(3, args.steps // 4))
    er = (out["P"].double() - o64["P"].double()).abs() / o64["P"].double()
    # the covariance on the matrix cores (BF16 hi/lo split) per row run, then the FP32 solves
    # from it (FP64 fallback as before)
    ms_mf = _timed(torch, lambda: step(mf=True), max(3, args.steps // 4))
    em = (out["P"].double() - o64["P"].double()).abs() / o64["P"].double()
    line = _base(args, "Capon pixels/sec, expanding aperture M(z)=16..128, L(z)=M(z)/2, FP32 + FP64 fallback",
                 npx / (ms * 1e-3), "pixels/s", ms,
                 {"workload": "expanding-aperture frame", "Nx": Nx, "Nz": Nz, "Mmax": Mmax, "Yavg": Y,
                  "pixels": npx, "ill_rows_d": 1e-5, "fallback_pixels": nfb})
    line["roofline"] = {"bound": "valu", "achieved": round(fpx * npx / (ms * 1e-3) / 1e12, 3),
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(fpx * npx / (ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                        "flop_per_pixel_mean": round(fpx)}
    line["all_fp64_ms_per_frame"] = round(ms64, 3)
    line["speedup_vs_all_fp64"] = round(ms64 / ms, 2)
    line["accuracy"] = {"vs_all_fp64_P_rel_err_max": float(e.max()),
                        "vs_all_fp64_P_rel_err_p999": float(torch.quantile(e[::97].float(), 0.999)),
                        "vs_all_fp64_P_rel_err_mean": float(e.mean()),
                        "vs_oracle_P_rel_err_max_64px": float(eo.max()),
                        "vs_oracle_w_rel_err_max_64px": float(max(ew))}
    # the library's untuned rule: the 16-lane shape with the fewest columns that holds L
    # (1, 2 or 3 columns: shapes 1, 2, 4), else 64 lanes (shape 3)
    default = {l: (1 if l <= 16 else 2 if l <= 32 else 4 if l <= 48 else 3) for l in tuned["shapes"]}
    line["autotune"] = {"default_shapes_ms_per_frame": round(ms_default, 3), "tuning_s": round(tune_s, 2),
                        "lengths": len(tuned["shapes"]),
                        "changed": {l: s for l, s in tuned["shapes"].items() if s != default[l]}}
    line["refined"] = {"ms_per_frame": round(ms_ref, 3), "vs_all_fp64_P_rel_err_max": float(er.max()),
                       "vs_all_fp64_P_rel_err_mean": float(er.mean())}
    line["mfma_covariance"] = {"ms_per_frame": round(ms_mf, 3), "vs_all_fp64_P_rel_err_max": float(em.max()),
                               "vs_all_fp64_P_rel_err_mean": float(em.mean()),
                               "path": "BF16 hi/lo MFMA covariance per row run -> R in HBM -> FP32 solves"}
    if args.no_cpu:
        return line
    # CPU leg: the oracle's C FP64 port over every output row of 8 beams, one call per run of
    # rows with the same (L, M, first element, d); pixels/s of that subset
    from oracle import c_capon
    xs8 = x[:8].cpu().numpy()
    Nzp = Nz - 2 * Y
    c = np.arange(Nzp) + Y
    key = np.stack([Lz[c], Mz[c], off[c], dz[c].view(np.int32)], axis=1)
    cores, _ = _host()
    t0 = time.perf_counter()
    jj = 0
    while jj < Nzp:
        j2 = jj + 1
        while j2 < Nzp and (key[j2] == key[jj]).all():
            j2 += 1
        Lr, Mr, orr = int(Lz[c[jj]]), int(Mz[c[jj]]), int(off[c[jj]])
        c_capon(xs8[:, :, orr:orr + Mr], float(dz[c[jj]]), Lr, Y, j=(jj, j2), want_w=True, nthreads=cores)
        jj = j2
    tc = time.perf_counter() - t0
    line["cpu_baseline"] = _cpu_line(8 * Nzp / tc, "pixels/s",
                                     f"8 of {Nx} beams x all {Nzp} rows (FP64 C port, OpenMP over pixels) in "
                                     f"{tc:.1f} s; the frame's rate is the same per pixel")
    line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def beamspace(args):
    import torch
    from cos_amd import CaponPlan, synth
    from oracle import capon_oracle as O
    Nx, Nz, M, L, Y, Nb, d = 512, 8192, 96, 48, 4, 5, 0.01
    x = synth.config4_torch(Nx, Nz, M, seed=4, device="cuda")
    plan = CaponPlan()
    npx = Nx * (Nz - 2 * Y)
    out = {"P": torch.empty(npx, device="cuda")}
    step = lambda: plan.run_beamspace(x, d, L, Y, Nb, want=("P",), out=out)
    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    K = M - L + 1
    fpx = 8 * Nb * L + 10 * Nb * (K - 1) + 8 * Nb * Nb * K + 4 * Nb * Nb + (4 * Nb ** 3) // 3 + 8 * Nb * Nb + 32 * Nb
    bpx = 8 * M + 4
    rng = np.random.default_rng(4)
    ix, jj = rng.integers(0, Nx, 128), rng.integers(0, Nz - 2 * Y, 128)
    win = np.stack([x[i, j:j + 2 * Y + 1, :].cpu().numpy() for i, j in zip(ix, jj)])
    ref = O.capon_beamspace(win.astype(np.complex128), d, L, Y, Nb,
                            pixels=(np.arange(len(ix)), np.zeros(len(ix), int)))
    P = out["P"].cpu().numpy().reshape(Nx, Nz - 2 * Y)[ix, jj]
    e = np.abs(P - ref["P"]) / ref["P"]
    ow = plan.run_beamspace(x, d, L, Y, Nb, want=("w",))      # element-space weights, untimed
    W = ow["w"].cpu().numpy().reshape(Nx, Nz - 2 * Y, L)[ix, jj]
    Wr = ref["w"].reshape(len(ix), L)
    ew = np.linalg.norm(W - Wr, axis=-1) / np.linalg.norm(Wr, axis=-1)
    gbs = bpx * npx / (ms * 1e-3) / 1e9
    line = _base(args, "Capon pixels/sec, sonar beamspace M=96 L=48 Nb=5 (sliding DFT)", npx / (ms * 1e-3),
                 "pixels/s", ms, {"workload": "sonar beamspace frame", "Nx": Nx, "Nz": Nz, "M": M, "L": L,
                                  "Nb": Nb, "Yavg": Y, "pixels": npx})
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_pixel": bpx,
                        "fp32_tflops": round(fpx * npx / (ms * 1e-3) / 1e12, 2), "flop_per_pixel": fpx}
    line["accuracy"] = {"P_rel_err_max_128px": float(e.max()), "w_rel_err_max_128px": float(ew.max())}
    if args.no_cpu:
        return line
    # CPU leg: the numpy FP64 restatement (sliding DFT, Nb x Nb Gram, Cholesky), vectorised
    # over pixels, on 2 beams x 2048 rows
    xc = x[:2, :2048 + 2 * Y].cpu().numpy().astype(np.complex128)
    t0 = time.perf_counter()
    O.capon_beamspace(xc, d, L, Y, Nb)
    tc = time.perf_counter() - t0
    line["cpu_baseline"] = _cpu_line(2 * 2048 / tc, "pixels/s",
                                     f"2 of {Nx} beams x 2048 rows, numpy FP64 restatement in {tc:.1f} s")
    line["cpu_baseline"]["cores"] = 1
    line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def steering(args):
    import torch
    from cos_amd import CaponPlan, synth
    from oracle import capon_oracle as O
    L, Y, d = 16, 2, 0.01
    xn = synth.config1(N=1024, M=32, seed=1)
    sines = synth.steering_sines(128)
    x = torch.from_numpy(xn).cuda()
    st = torch.from_numpy(sines.astype(np.float32)).cuda()
    plan = CaponPlan()
    ref = O.capon_steered(xn.astype(np.complex128), d, L, Y, sines.astype(np.float32).astype(np.float64))

    def leg(refine):
        step = lambda: plan.run(x, d, L, Y, steer_sin=st, want=("P", "w"), refine=refine)
        for _ in range(args.warmup):
            step()
        ms = _timed(torch, step, args.steps)
        o = step()
        torch.cuda.synchronize()
        P = o["P"].cpu().numpy().reshape(-1, 128)
        e = np.abs(P - ref["P"]) / ref["P"]
        W = o["w"].cpu().numpy().reshape(-1, 128, L)
        Wr = ref["w"].reshape(W.shape)
        ew = np.linalg.norm(W - Wr, axis=-1) / np.linalg.norm(Wr, axis=-1)
        return ms, P, e, ew

    # one refinement step keeps every power within 1e-4 of FP64 here (the strong sources
    # take the unrefined FP32 solve to ~1.3e-4 on a few bins); the unrefined rate beside it
    ms0, _, e0, ew0 = leg(False)
    ms, P, e, ew = leg(True)
    npow = P.size
    line = _base(args, "Capon steered powers/sec, ULA M=32 L=16, 128 angles per range bin", npow / (ms * 1e-3),
                 "powers/s", ms, {"workload": "far-field steered ULA", "bins": 1020, "angles": 128, "M": 32, "L": L,
                                  "Yavg": Y, "factorisations": 1020, "precision": "fp32 + one refinement step"})
    line["accuracy"] = {"P_rel_err_max": float(e.max()), "w_rel_err_max": float(ew.max())}
    line["unrefined"] = {"ms": round(ms0, 4), "powers_per_s": npow / (ms0 * 1e-3), "P_rel_err_max": float(e0.max()),
                         "w_rel_err_max": float(ew0.max())}
    if args.no_cpu:
        return line
    # CPU leg: the reference's way to steer, one drop-in call per angle on the cube
    # phase-rotated by exp(-j pi m sin theta), through the oracle's C FP64 port; all 128 angles
    from oracle import c_capon
    cores, _ = _host()
    m = np.arange(xn.shape[-1])
    t0 = time.perf_counter()
    for sn in sines.astype(np.float32).astype(np.float64):
        c_capon((xn * np.exp(-1j * np.pi * m * sn)).astype(np.complex64), d, L, Y, want_w=True, nthreads=cores)
    tc = time.perf_counter() - t0
    line["cpu_baseline"] = _cpu_line(npow / tc, "powers/s",
                                     f"all 128 angles x 1020 bins, one FP64 C-port call per rotated cube, {tc:.2f} s")
    line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def shard(args):
    """One GPU's share of the headline frame split by beam over 8 GPUs: 32 beams x 2048
    ranges x 64 elements, L = 32 (pixels need no exchange, so this is a whole shard)."""
    import torch
    from cos_amd import CaponPlan, synth
    G, Nx, Nz, M, L, Y, d = 8, 256, 2048, 64, 32, 2, 0.01
    xn = synth.config2(Nx, Nz, M, seed=2)[: Nx // G].copy()
    x = torch.from_numpy(xn).cuda()
    plan = CaponPlan()
    npx = (Nx // G) * (Nz - 2 * Y)
    out = {"P": torch.empty(npx, device="cuda"), "z": torch.empty(npx, dtype=torch.complex64, device="cuda"),
           "w": torch.empty((npx, L), dtype=torch.complex64, device="cuda")}
    step = lambda: plan.run(x, d, L, Y, precision="fp32+fallback", out=out, graph=True)
    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    line = _base(args, "Capon pixels/sec on one beam shard (1/8 of the 256x2048x64 L=32 frame)",
                 npx / (ms * 1e-3), "pixels/s", ms,
                 {"workload": "beam shard of the per-pixel frame", "shards": G, "Nx": Nx // G, "Nz": Nz, "M": M,
                  "L": L, "Yavg": Y, "pixels": npx})
    line["note"] = ("the shard needs no halo or exchange (pixels are independent along beams); "
                    "G GPUs would each run this launch on their own beams")
    if not args.no_cpu:   # the oracle's C FP64 port over the whole shard, all host threads
        from oracle import c_capon
        cores, _ = _host()
        t0, reps = time.perf_counter(), 0
        while reps < 1 or time.perf_counter() - t0 < 1.0:
            c_capon(xn, d, L, Y, want_w=True, nthreads=cores)
            reps += 1
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = _cpu_line(reps * npx / tc, "pixels/s",
                                         f"the whole shard ({Nx // G} beams x {Nz - 2 * Y} rows) {reps} times, FP64 "
                                         f"C port, OpenMP over pixels, {tc:.2f} s")
        line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def adjoint(args):
    import torch
    from cos_amd import synth
    from cos_amd.autograd import capon_power
    F, Nx, Nz, M, L, Y, d = 8, 128, 512, 64, 32, 2, 0.01
    frames = torch.stack([synth.speckle_cube_torch(Nx, Nz, M, seed=50 + f, device="cuda") for f in range(F)])
    phi = torch.from_numpy(synth.config5_phase(M, seed=5)).cuda().requires_grad_(True)
    npx = F * Nx * (Nz - 2 * Y)

    def fwd():
        xs = (frames * torch.exp(1j * phi)).reshape(F * Nx, Nz, M)
        return -torch.log(capon_power(xs, d, L, Y)).mean()

    def step():
        phi.grad = None
        fwd().backward()

    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    with torch.no_grad():
        ms_f = _timed(torch, lambda: fwd(), args.steps)
    line = _base(args, "Capon pixels/sec through forward + hand-written adjoint (dLoss/dphi, dLoss/dx)",
                 npx / (ms * 1e-3), "pixels/s", ms,
                 {"workload": "differentiable frames", "frames": F, "Nx": Nx, "Nz": Nz, "M": M, "L": L,
                  "Yavg": Y, "pixels": npx})
    line["forward_ms"] = round(ms_f, 3)
    line["backward_over_forward"] = round((ms - ms_f) / ms_f, 2)
    # accuracy at this size: dLoss/dphi of the adjoint against central differences
    # (h = 1e-3) of the same loss through the FP64 kernel, on 8 seeded coordinates; and
    # the forward power of frame 0 against the FP64 oracle on 2 beams x 32 rows
    from cos_amd import CaponPlan
    from oracle import c_capon
    phi.grad = None
    fwd().backward()
    g = phi.grad.detach().double().cpu().numpy()
    plan64 = CaponPlan()

    def loss64(ph):   # phases in FP64, the cube rounded to complex64 as the kernels read it
        xs = (frames.to(torch.complex128) * torch.exp(1j * ph)).to(torch.complex64)
        P = plan64.run(xs.reshape(F * Nx, Nz, M).contiguous(), d, L, Y, precision="fp64", want=("P",))["P"]
        return float(-torch.log(P.double()).mean())

    h = 1e-3
    coords = np.random.default_rng(6).choice(M, 8, replace=False)
    phi0 = phi.detach().double()
    fd = []
    with torch.no_grad():
        for m in coords:
            e = torch.zeros(M, dtype=torch.float64, device="cuda")
            e[m] = h
            fd.append((loss64(phi0 + e) - loss64(phi0 - e)) / (2 * h))
    fd = np.array(fd)
    gerr = np.abs(g[coords] - fd) / np.abs(g).max()    # against the gradient's scale
    with torch.no_grad():
        x0 = (frames[0] * torch.exp(1j * phi)).contiguous()
        P0 = capon_power(x0, d, L, Y).reshape(Nx, Nz - 2 * Y)[:2, 100:132].double().cpu().numpy()
    ref = c_capon(x0.cpu().numpy(), d, L, Y, ix=(0, 2), j=(100, 132), want_w=False)
    eP = np.abs(P0 - ref["P"]) / ref["P"]
    line["accuracy"] = {"dphi_err_vs_central_differences_max": float(gerr.max()),
                        "dphi_sample": "8 seeded element phases, error over max |dLoss/dphi|",
                        "P_rel_err_max_64px": float(eP.max())}
    if args.no_cpu:
        return line
    # CPU leg: the reference has no gradient; the oracle's C FP64 port runs the forward power
    # of all 8 phase-corrected frames (the reference's own work per frame)
    cores, _ = _host()
    with torch.no_grad():
        xh = (frames * torch.exp(1j * phi)).reshape(F * Nx, Nz, M).cpu().numpy()
    t0 = time.perf_counter()
    c_capon(xh, d, L, Y, want_w=False, nthreads=cores)
    tc = time.perf_counter() - t0
    line["cpu_baseline"] = _cpu_line(npx / tc, "pixels/s",
                                     f"forward only (no reference gradient exists): all {npx} pixels of the 8 frames, "
                                     f"FP64 C port, {tc:.1f} s")
    line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def tdib(args):
    """Delay stage producing the headline frame's cube: 256 beams x 2048 range pixels x 64
    hydrophones from 64 x 32768 matched-filtered samples; cube and image written."""
    import torch
    from cos_amd import stages
    n_h, n_t, n_x, n_y = 64, 32768, 256, 2048
    g = torch.Generator(device="cuda").manual_seed(7)
    mf = torch.complex(torch.randn(n_h, n_t, device="cuda", generator=g),
                       torch.randn(n_h, n_t, device="cuda", generator=g))
    mtaxe = 0.01 + np.arange(n_t) * 4.0e-6
    x_r = torch.linspace(0.0, 1.5, n_h, dtype=torch.float64, device="cuda")
    D = torch.full((n_h,), 0.3, dtype=torch.float64, device="cuda")
    xarr = torch.linspace(-6.0, 6.0, n_x, dtype=torch.float64, device="cuda")[:, None].expand(n_x, n_y).contiguous()
    yarr = torch.linspace(15.0, 90.0, n_y, dtype=torch.float64, device="cuda")[None, :].expand(n_x, n_y).contiguous()
    step = lambda: stages.tdib_cube(mf, mtaxe, x_r, D, 1500.0, 1.0e5, xarr, yarr, fill=0.0)
    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    npx = n_x * n_y
    # bytes: cube written (8 n_h per pixel) + image (8) + coordinates read (16); channel
    # samples are read from cache (each pixel reads 2 samples per hydrophone)
    bpx = 8 * n_h + 8 + 16
    gbs = bpx * npx / (ms * 1e-3) / 1e9
    line = _base(args, "delay stage: pixels/sec producing the 256x2048x64 Capon cube", npx / (ms * 1e-3),
                 "pixels/s", ms, {"workload": "TDIB delay stage", "n_x": n_x, "n_y": n_y, "n_h": n_h, "n_t": n_t},
                 dtype="f64 geometry, f32 data")
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_pixel": bpx}
    if not args.no_cpu:   # the oracle's numpy FP64 restatement (np.interp per hydrophone) on 64 beams
        from oracle import stages_oracle
        nb = 64
        mfn, xrn, Dn = mf.cpu().numpy(), x_r.cpu().numpy(), D.cpu().numpy()
        xan, yan = xarr[:nb].cpu().numpy(), yarr[:nb].cpu().numpy()
        t0 = time.perf_counter()
        stages_oracle.tdib_cube(mfn, mtaxe, xrn, Dn, 1500.0, 1.0e5, xan, yan)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = _cpu_line(nb * n_y / tc, "pixels/s",
                                         f"{nb} of {n_x} beams x {n_y} pixels x {n_h} hydrophones, numpy FP64 "
                                         f"restatement in {tc:.2f} s", extra={"cores": 1})
        line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def lca(args):
    """Window selection over 8 windows on a 256 x 2048 image, 5 x 1 range filter."""
    import torch
    from cos_amd import stages
    Nx, Ny, Nw = 256, 2048, 8
    g = torch.Generator(device="cuda").manual_seed(8)
    Z = torch.complex(torch.randn(Nx, Ny, Nw, device="cuda", generator=g),
                      torch.randn(Nx, Ny, Nw, device="cuda", generator=g))
    wrf = torch.ones((5, 1), dtype=torch.float32, device="cuda")
    # reference mode (lcacx.pyx's own output: FP64 sliding sums per beam, the filter read to
    # the host each call) is the drop-in; the centred mode is timed beside it
    step = lambda: stages.lca(Z, wrf, mode="reference")
    for _ in range(args.warmup):
        step()
    ms = _timed(torch, step, args.steps)
    cstep = lambda: stages.lca(Z, wrf, mode="centred")
    for _ in range(args.warmup):
        cstep()
    ms_c = _timed(torch, cstep, args.steps)
    npx = Nx * Ny
    bpx = 8 * Nw + 8 + 4
    gbs = bpx * npx / (ms * 1e-3) / 1e9
    line = _base(args, "LCA window selection: pixels/sec, 8 windows, 5x1 filter", npx / (ms * 1e-3), "pixels/s", ms,
                 {"workload": "LCA window selection", "Nx": Nx, "Ny": Ny, "Nw": Nw, "filter": [5, 1],
                  "mode": "reference (lcacx.pyx output)"}, dtype="f64")
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_pixel": bpx}
    gbs_c = bpx * npx / (ms_c * 1e-3) / 1e9
    line["centred"] = {"ms_per_frame": round(ms_c, 4), "pixels_per_s": npx / (ms_c * 1e-3),
                       "GB_s": round(gbs_c, 1), "frac": round(gbs_c / PEAK_HBM_GBS, 4)}
    if not args.no_cpu:   # the oracle's reference-mode restatement (numpy, vectorised over beams)
        from oracle import stages_oracle
        Zn, wn = Z.cpu().numpy(), wrf.cpu().numpy()
        t0 = time.perf_counter()
        stages_oracle.lca_reference(Zn, wn)
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = _cpu_line(npx / tc, "pixels/s",
                                         f"the whole {Nx} x {Ny} x {Nw} frame, numpy FP64 restatement of the "
                                         f"reference's sliding sums in {tc:.2f} s", extra={"cores": 1})
        line["speedup_vs_cpu"] = round(line["value"] / line["cpu_baseline"]["value"], 1)
    return line


def solver(args):
    """Batched general solve (Solver::solve, device operands) at N = 32 over the headline
    frame's 523,264 systems, and N = 72 over 65,536."""
    import torch
    from cos_amd import Solver
    from oracle import solver_oracle as SO
    res = {}
    for N, batch in ((32, 256 * 2044), (72, 65536)):
        g = torch.Generator(device="cuda").manual_seed(N)
        A = torch.complex(torch.randn(batch, N, N, device="cuda", generator=g),
                          torch.randn(batch, N, N, device="cuda", generator=g))
        A += 4 * torch.eye(N, device="cuda")
        b = torch.complex(torch.randn(batch, N, device="cuda", generator=g),
                          torch.randn(batch, N, device="cuda", generator=g))
        x = torch.empty_like(b)
        S = Solver()
        step = lambda: S.solve(x.data_ptr(), A.data_ptr(), b.data_ptr(), N, batch, x_on_gpu=True, A_on_gpu=True,
                               b_on_gpu=True)
        for _ in range(args.warmup):
            assert step() == 0
        ms = _timed(torch, step, args.steps)
        # accuracy on a sample against the oracle (complex128)
        ns = 256
        Ah = A[:ns].cpu().numpy()
        ref = SO.solve_batch(Ah.reshape(ns, -1), b[:ns].cpu().numpy())
        err = float(np.abs(x[:ns].cpu().numpy() - ref).max() / np.abs(ref).max())
        # HBM floor: A, b read and x written once
        byt = 8 * (N * N + 2 * N) * batch
        flop = (8 * N ** 3) // 3 + 8 * N * N
        ncpu = 2048 if N == 32 else 512
        t0 = time.perf_counter()
        SO.solve_batch(A[:ncpu].cpu().numpy().reshape(ncpu, -1), b[:ncpu].cpu().numpy())
        tc = time.perf_counter() - t0
        res[N] = dict(batch=batch, ms=ms, sys_s=batch / (ms * 1e-3), gbs=byt / (ms * 1e-3) / 1e9,
                      tflops=flop * batch / (ms * 1e-3) / 1e12, err=err, cpu=ncpu / tc, ncpu=ncpu)
    r = res[32]
    line = _base(args, "Solver::solve: systems/sec, N = 32, device operands", r["sys_s"], "systems/s", r["ms"],
                 {"workload": "batched general solve", "N": 32, "batch": r["batch"],
                  "pivoting": "partial, |re|+|im|"})
    line["roofline"] = {"bound": "hbm", "achieved": round(r["gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(r["gbs"] / PEAK_HBM_GBS, 4), "bytes_per_system": 8 * (32 * 32 + 64),
                        "fp32_tflops": round(r["tflops"], 2),
                        "issue": "VALU-issue bound (rows in registers, 2 systems per wave)"}
    line["accuracy"] = {"x_rel_err_max": r["err"], "sample": "256 systems vs the oracle (complex128)"}
    r72 = res[72]
    line["N72"] = {"batch": r72["batch"], "ms": round(r72["ms"], 4), "systems_per_s": r72["sys_s"],
                   "GB_s": round(r72["gbs"], 1), "fp32_tflops": round(r72["tflops"], 2), "x_rel_err_max": r72["err"]}
    line["cpu_baseline"] = _cpu_line(r["cpu"], "systems/s", f"{r['ncpu']} systems of N = 32 through the oracle "
                                     "(numpy, complex128, vectorised over the batch)", extra={"cores": 1})
    return line
