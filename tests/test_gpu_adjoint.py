"""The hand-written adjoint (cos_capon_backward) against an FP64 autograd reference and
against central finite differences of the FP64 forward kernel (GPU).

The reference gradient is torch autograd on the CPU in complex128 through
torch.linalg.solve -- a library routine independent of the kernel's Cholesky; the
tolerances are written in each test (1e-3, the differentiable workload's bound)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def torch_capon_power(x, d, L, Y):
    """P for every pixel of x [Nx][Nz][M] (complex128 CPU tensor), differentiable."""
    import torch
    Nx, Nz, M = x.shape
    K, T = M - L + 1, 2 * Y + 1
    rows = x.unfold(1, T, 1).permute(0, 1, 3, 2)                  # [Nx][Nzp][T][M]
    idx = torch.arange(K)[:, None] + torch.arange(L)[None, :]
    u = rows[..., idx]                                             # [Nx][Nzp][T][K][L]
    S = torch.einsum("abtkr,abtkc->abrc", u, u.conj())
    tr = torch.diagonal(S, dim1=-2, dim2=-1).real.sum(-1)
    SL = S + (d * tr / L)[..., None, None] * torch.eye(L, dtype=S.dtype)
    ones = torch.ones(L, dtype=S.dtype)
    y = torch.linalg.solve(SL, ones.expand(SL.shape[:-1]).unsqueeze(-1)).squeeze(-1)
    q = y.sum(-1).real
    return (1.0 / (K * T * q)).reshape(-1)


def fp64_power(x_c64, d, L, Y):
    import torch
    from cos_amd import CaponPlan
    P = CaponPlan().run(x_c64, d, L, Y, precision="fp64", want=("P",))["P"]
    torch.cuda.synchronize()
    return P.double()


@pytest.mark.parametrize("M,L,Y,d", [(16, 8, 1, 0.05), (64, 32, 2, 0.01), (24, 13, 0, 0.2)])
def test_adjoint_matches_fp64_autograd(M, L, Y, d):
    import torch
    from cos_amd.autograd import capon_power
    say("adjoint vs autograd", M, L, Y, d)
    rng = np.random.default_rng(M + L)
    xn = (rng.standard_normal((3, 12, M)) + 1j * rng.standard_normal((3, 12, M))).astype(np.complex64)
    wts = torch.from_numpy(rng.uniform(0.5, 1.5, 3 * (12 - 2 * Y)))
    # reference
    xr = torch.from_numpy(xn.astype(np.complex128)).requires_grad_(True)
    dr = torch.tensor(d, dtype=torch.float64, requires_grad=True)
    lr = -(wts * torch.log(torch_capon_power(xr, dr, L, Y))).mean()
    lr.backward()
    # GPU adjoint
    xg = torch.from_numpy(xn).cuda().requires_grad_(True)
    dg = torch.tensor(d, dtype=torch.float32, device="cuda", requires_grad=True)
    Pg = capon_power(xg, dg, L, Y)
    lg = -(wts.cuda().float() * torch.log(Pg)).mean()
    lg.backward()
    gx = xg.grad.cpu().numpy()
    gref = xr.grad.numpy()
    err = np.abs(gx - gref).max() / np.abs(gref).max()
    print(f"grad x rel err {err:.2e}; grad d {dg.grad.item():.6e} vs {dr.grad.item():.6e}", flush=True)
    assert err <= 1e-3
    assert abs(dg.grad.item() - dr.grad.item()) <= 1e-3 * abs(dr.grad.item())
    assert abs(lg.item() - lr.item()) <= 1e-4 * abs(lr.item())


def test_adjoint_central_differences():
    """32 seeded coordinates of x (real and imaginary parts) and the loading, h = 1e-3,
    against the FP64 forward kernel."""
    import torch
    from cos_amd import synth
    from cos_amd.autograd import capon_power
    say("adjoint central differences")
    Nx, Nz, M, L, Y, d = 4, 40, 64, 32, 2, 0.01
    xn = synth.speckle_cube(Nx, Nz, M, seed=61)
    xg = torch.from_numpy(xn).cuda().requires_grad_(True)
    dg = torch.tensor(d, device="cuda", requires_grad=True)
    loss = -torch.log(capon_power(xg, dg, L, Y)).mean()
    loss.backward()
    g = xg.grad.cpu().numpy()
    rng = np.random.default_rng(6)
    h = 1e-3

    def f(xx, dd=d):
        return float(-torch.log(fp64_power(torch.from_numpy(xx).cuda(), dd, L, Y)).mean())

    worst = 0.0
    for n in range(32):
        ix, iz, m = rng.integers(Nx), rng.integers(Nz), rng.integers(M)
        part = n % 2
        e = np.zeros_like(xn)
        e[ix, iz, m] = h if part == 0 else 1j * h
        fd = (f(xn + e) - f(xn - e)) / (2 * h)
        an = g[ix, iz, m].real if part == 0 else g[ix, iz, m].imag
        scale = np.abs(g).max()
        worst = max(worst, abs(fd - an) / scale)
    fd_d = (f(xn, d * (1 + h)) - f(xn, d * (1 - h))) / (2 * d * h)
    print(f"FD worst rel err (x) {worst:.2e}; d: fd {fd_d:.6e} adjoint {dg.grad.item():.6e}", flush=True)
    assert worst <= 1e-3
    assert abs(fd_d - dg.grad.item()) <= 1e-3 * abs(fd_d)


def test_phase_gradient_differentiable_workload():
    """Per-element phase errors phi on seeded frames, loss = -mean log P; dLoss/dphi from
    the adjoint chained by torch, against central differences of the FP64 forward."""
    import torch
    from cos_amd import synth
    from cos_amd.autograd import capon_power
    say("phase gradient")
    F, Nx, Nz, M, L, Y, d = 2, 16, 96, 64, 32, 2, 0.01
    frames = torch.from_numpy(np.stack([synth.speckle_cube(Nx, Nz, M, seed=50 + f) for f in range(F)])).cuda()
    phi0 = torch.from_numpy(synth.config5_phase(M, seed=5)).double().cuda()
    phi = phi0.clone().float().requires_grad_(True)
    xs = (frames * torch.exp(1j * phi)).reshape(F * Nx, Nz, M)
    loss = -torch.log(capon_power(xs, d, L, Y)).mean()
    loss.backward()
    g = phi.grad.double().cpu().numpy()

    def f(ph):
        xx = (frames.to(torch.complex128) * torch.exp(1j * ph)).to(torch.complex64).reshape(F * Nx, Nz, M)
        return float(-torch.log(fp64_power(xx.contiguous(), d, L, Y)).mean())

    h = 1e-3
    rng = np.random.default_rng(6)
    for m in rng.choice(M, 8, replace=False):
        e = torch.zeros(M, dtype=torch.float64, device="cuda")
        e[m] = h
        fd = (f(phi0 + e) - f(phi0 - e)) / (2 * h)
        print(f"phi[{m}] fd {fd:.6e} adjoint {g[m]:.6e}", flush=True)
        assert abs(fd - g[m]) <= 1e-3 * np.abs(g).max() + 1e-6


def test_differentiable_workload_full_size():
    """The differentiable workload at its stated size: 8 seeded frames of 128 beams x 512
    ranges x 64 elements, L = 32, Yavg = 2, per-element phase errors phi (seeded, in
    [-0.5, 0.5] rad), loss = -mean log P over the 520,192 pixels.  The adjoint's dLoss/dphi,
    dLoss/d(snapshots) and dLoss/dd against central differences (h = 1e-3) of the FP64
    forward kernel on 32 seeded coordinates (16 of phi, 16 of the snapshots) and d, each
    within 1e-3 of the largest gradient entry of its kind (of |dLoss/dd| for d)."""
    import torch
    from cos_amd import synth
    from cos_amd.autograd import capon_power
    F, Nx, Nz, M, L, Y, d = 8, 128, 512, 64, 32, 2, 0.01
    say("differentiable workload, full size", F, Nx, Nz, M, L, Y)
    frames0 = torch.stack([synth.speckle_cube_torch(Nx, Nz, M, seed=50 + f) for f in range(F)])
    phi0 = torch.from_numpy(synth.config5_phase(M, seed=5)).double().cuda()
    frames = frames0.clone().requires_grad_(True)
    phi = phi0.clone().float().requires_grad_(True)
    dg = torch.tensor(d, device="cuda", requires_grad=True)
    xs = (frames * torch.exp(1j * phi)).reshape(F * Nx, Nz, M)
    P = capon_power(xs, dg, L, Y)
    assert P.numel() == F * Nx * (Nz - 2 * Y) == 520192
    loss = -torch.log(P).mean()
    loss.backward()
    gphi = phi.grad.double().cpu().numpy()
    gx = frames.grad.cpu().numpy()
    gd = dg.grad.item()
    f64 = frames0.to(torch.complex128)

    def f(ph=phi0, dd=d, pert=None):
        xx = f64 * torch.exp(1j * ph)
        if pert is not None:
            xx = xx.clone()
            xx[pert[0]] += pert[1]
        xx = xx.to(torch.complex64).reshape(F * Nx, Nz, M).contiguous()
        return float(-torch.log(fp64_power(xx, dd, L, Y)).mean())

    h = 1e-3
    rng = np.random.default_rng(6)
    worst_phi = 0.0
    for m in rng.choice(M, 16, replace=False):
        e = torch.zeros(M, dtype=torch.float64, device="cuda")
        e[m] = h
        fd = (f(phi0 + e) - f(phi0 - e)) / (2 * h)
        worst_phi = max(worst_phi, abs(fd - gphi[m]) / np.abs(gphi).max())
    # snapshot coordinates: dLoss/dRe and dLoss/dIm of x, where x is the frame before the
    # phase rotation (the chain through exp(j phi) is torch's)
    worst_x = 0.0
    scale_x = np.abs(gx).max()
    for n in range(16):
        fi, ix, iz, m = int(rng.integers(F)), int(rng.integers(Nx)), int(rng.integers(Y, Nz - Y)), int(rng.integers(M))
        step = h if n % 2 == 0 else 1j * h
        rot = torch.exp(1j * phi0[m])
        fp = f(pert=((fi, ix, iz, m), step * rot))
        fm = f(pert=((fi, ix, iz, m), -step * rot))
        fd = (fp - fm) / (2 * h)
        an = gx[fi, ix, iz, m].real if n % 2 == 0 else gx[fi, ix, iz, m].imag
        worst_x = max(worst_x, abs(fd - an) / scale_x)
    fd_d = (f(dd=d * (1 + h)) - f(dd=d * (1 - h))) / (2 * d * h)
    print(f"loss {loss.item():.6f}; FD worst rel err phi {worst_phi:.2e}, x {worst_x:.2e}; "
          f"d: fd {fd_d:.6e} adjoint {gd:.6e}", flush=True)
    assert abs(loss.item() - f()) <= 1e-5 * abs(loss.item())
    assert worst_phi <= 1e-3
    assert worst_x <= 1e-3
    assert abs(fd_d - gd) <= 1e-3 * abs(fd_d)
