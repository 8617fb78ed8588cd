"""Seeded synthetic inputs for the Capon workloads.

Every random number comes from a counter-based generator: SplitMix64 (the
finaliser of Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) applied to ``key + (counter + 1) * GOLDEN``, the top
53 bits scaled to [0, 1), and the Box-Muller transform for Gaussians.  The
stream depends only on (seed, stream id, counter), so any slice of a cube can be
regenerated independently and the same numbers can be produced by other
implementations of the same recipe.

Cube layout is the reference's ``x[Nx][Nz][M]`` complex64, M fastest
(framework/beamformer/capon/RealTimeCapon/buildR_kernel.cu:18,86;
framework/beamformer/capon/getCaponCUDA.pyx:20-22).
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def _key(seed: int, stream: int) -> np.uint64:
    k = np.array([(seed * 0x100000001B3 + stream * 0x1000193 + 0x632BE59BD9B4E019) & (2**64 - 1)],
                 dtype=np.uint64)
    return _mix(k)[0]


def uniforms(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n uniforms in [0, 1) from counters offset..offset+n-1 (float64)."""
    with np.errstate(over="ignore"):
        ctr = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = _mix(_key(seed, stream) + ctr * GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cnormal(seed: int, stream: int, shape, chunk: int = 1 << 22) -> np.ndarray:
    """Unit-variance circular complex Gaussian (real and imaginary parts N(0, 1/2)),
    complex128, generated in chunks to bound temporaries."""
    n = int(np.prod(shape))
    out = np.empty(n, dtype=np.complex128)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        u = uniforms(seed, stream, 2 * (e - s), offset=2 * s)
        r = np.sqrt(-np.log1p(-u[0::2]))          # sqrt(-2 ln(1-u1)) * sqrt(1/2)
        th = 2.0 * np.pi * u[1::2]
        out[s:e] = r * np.cos(th) + 1j * r * np.sin(th)
    return out.reshape(shape)


AXIAL_PULSE = np.array([0.25, 0.5, 1.0, 0.5, 0.25])


def speckle_cube(Nx: int, Nz: int, M: int, seed: int, n_scatterers: int = 10,
                 scatterer_db: float = 30.0, stream: int = 0) -> np.ndarray:
    """Pre-delayed near-field cube: axially filtered circular Gaussian speckle plus
    point scatterers that are coherent (identical) across all M elements.

    Speckle: CN(0,1) filtered along z by the 5-tap pulse [0.25,0.5,1,0.5,0.25]
    (variance 1.625).  Each scatterer sits at a seeded uniform (z, x) cell with
    power 10^(dB/10) times the speckle variance and a seeded phase, and passes
    through the same axial pulse.  Returns complex64 [Nx][Nz][M].
    """
    n = cnormal(seed, stream, (Nx, Nz + 4, M))
    h = AXIAL_PULSE
    cube = np.zeros((Nx, Nz, M), dtype=np.complex128)
    for i in range(5):
        cube += h[i] * n[:, i:i + Nz, :]
    del n
    var = float(np.sum(h * h))
    amp = np.sqrt(10.0 ** (scatterer_db / 10.0) * var)
    for zs, xs, ph in _scatterers(Nx, Nz, seed, n_scatterers, stream):
        a = amp * np.exp(2j * np.pi * ph)
        for i in range(5):
            zz = zs + i - 2
            if 0 <= zz < Nz:
                cube[xs, zz, :] += h[i] * a
    return cube.astype(np.complex64)


def _scatterers(Nx, Nz, seed, n_scatterers, stream):
    u = uniforms(seed, stream + 1, 3 * n_scatterers)
    return [(min(int(u[3 * s] * Nz), Nz - 1), min(int(u[3 * s + 1] * Nx), Nx - 1), u[3 * s + 2])
            for s in range(n_scatterers)]


def scatterer_positions(Nx: int, Nz: int, seed: int, n_scatterers: int = 10, stream: int = 0):
    """(range row z, beam x) of each point scatterer speckle_cube() places."""
    return [(z, x) for z, x, _ in _scatterers(Nx, Nz, seed, n_scatterers, stream)]


def config2(Nx: int = 256, Nz: int = 2048, M: int = 64, seed: int = 2) -> np.ndarray:
    """Per-pixel near-field Capon workload: [Nx][Nz][M] speckle + 10 scatterers at 30 dB."""
    return speckle_cube(Nx, Nz, M, seed)


def config1(N: int = 1024, M: int = 32, seed: int = 1, snr_db: float = 20.0,
            angles_deg=(-20.0, 0.0, 15.0), centres=(200, 500, 800), half_span: int = 3):
    """Far-field ULA (half-wavelength spacing) with three sources present only in range
    bins centre +/- half_span.  Returns x as a [1][N][M] complex64 cube."""
    m = np.arange(M)
    x = cnormal(seed, 0, (N, M))
    amp = np.sqrt(10.0 ** (snr_db / 10.0))
    for k, (th, c) in enumerate(zip(angles_deg, centres)):
        s = cnormal(seed, 1 + k, (N,)) * amp
        mask = np.zeros(N)
        mask[max(0, c - half_span):min(N, c + half_span + 1)] = 1.0
        steer = np.exp(1j * np.pi * m * np.sin(np.deg2rad(th)))
        x += (mask * s)[:, None] * steer[None, :]
    return x[None].astype(np.complex64)


def steering_sines(n: int = 128, lo_deg: float = -45.0, hi_deg: float = 45.0) -> np.ndarray:
    return np.sin(np.deg2rad(np.linspace(lo_deg, hi_deg, n)))


def config3_geometry(Nz: int = 4096, seed: int = 3, ill_fraction: float = 0.05,
                     d: float = 0.01, d_ill: float = 1e-5, Mmax: int = 128):
    """Expanding aperture (f-number 1.5): active elements M(z) grow linearly from 16 to
    Mmax over the depth, L(z) = M(z)/2, centred in the Mmax-wide array.  5% of rows
    (seeded) use the small loading d_ill.  Returns (M_rows, L_rows, off_rows, d_rows)."""
    z = np.arange(Nz)
    Mz = 2 * np.round((16 + (Mmax - 16) * z / (Nz - 1)) / 2).astype(np.int64)
    Lz = Mz // 2
    off = (Mmax - Mz) // 2
    u = uniforms(seed, 7, Nz)
    dz = np.where(u < ill_fraction, d_ill, d).astype(np.float32)
    return Mz.astype(np.int32), Lz.astype(np.int32), off.astype(np.int32), dz


def config3(Nx: int = 256, Nz: int = 4096, seed: int = 3, Mmax: int = 128):
    """Speckle cube for the expanding aperture; elements outside the active aperture
    are zeroed.  Returns (cube [Nx][Nz][Mmax], M_rows, L_rows, off_rows, d_rows)."""
    Mz, Lz, off, dz = config3_geometry(Nz, seed, Mmax=Mmax)
    cube = speckle_cube(Nx, Nz, Mmax, seed)
    m = np.arange(Mmax)
    active = (m[None, :] >= off[:, None]) & (m[None, :] < (off + Mz)[:, None])
    cube *= active[None, :, :]
    return cube, Mz, Lz, off, dz


def config4(Nx: int = 512, Nz: int = 8192, M: int = 96, seed: int = 4,
            n_sources: int = 4, span: int = 200):
    """Sonar beamspace workload: unit CN noise plus plane waves at seeded angles in
    [-30, 30] deg and SNR in [10, 25] dB, each over a seeded range span, pre-delayed per
    look direction (look direction ix steers to sin(theta_ix) uniform in [-0.5, 0.5])."""
    x = cnormal(seed, 0, (Nx, Nz, M)).astype(np.complex64)
    u = uniforms(seed, 1, 4 * n_sources)
    m = np.arange(M)
    looks = np.linspace(-0.5, 0.5, Nx)
    for s in range(n_sources):
        th = np.deg2rad(-30.0 + 60.0 * u[4 * s])
        snr = 10.0 + 15.0 * u[4 * s + 1]
        z0 = int(u[4 * s + 2] * (Nz - span))
        amp = np.sqrt(10.0 ** (snr / 10.0))
        sig = cnormal(seed, 10 + s, (span,)) * amp
        for ix in range(Nx):
            ph = np.exp(1j * np.pi * m * (np.sin(th) - looks[ix]))
            x[ix, z0:z0 + span, :] += (sig[:, None] * ph[None, :]).astype(np.complex64)
    return x


def config5_frames(Nx: int = 128, Nz: int = 512, M: int = 64, n_frames: int = 8,
                   seed0: int = 50):
    return np.stack([speckle_cube(Nx, Nz, M, seed0 + f) for f in range(n_frames)])


def config5_phase(M: int = 64, seed: int = 5) -> np.ndarray:
    return (uniforms(seed, 0, M) - 0.5).astype(np.float32)


# ------------------------------------------------------------------------------------
# The same generator on a torch device (used for cubes too large for numpy).  int64
# arithmetic wraps like uint64; right shifts are made logical by masking.
# ------------------------------------------------------------------------------------

def _t_mix(z):
    import torch
    m30 = (1 << 34) - 1
    m27 = (1 << 37) - 1
    m31 = (1 << 33) - 1
    c1 = int(np.array([0xBF58476D1CE4E5B9], dtype=np.uint64).view(np.int64)[0])
    c2 = int(np.array([0x94D049BB133111EB], dtype=np.uint64).view(np.int64)[0])
    z = z ^ ((z >> 30) & m30)
    z = z * c1
    z = z ^ ((z >> 27) & m27)
    z = z * c2
    return z ^ ((z >> 31) & m31)


def cnormal_torch(seed: int, stream: int, shape, device, chunk: int = 1 << 24, dtype=None):
    """Same numbers as cnormal() (complex128 arithmetic, returned as complex64) computed
    with torch on `device`, in chunks."""
    import torch
    n = int(np.prod(shape))
    key = int(np.array([_key(seed, stream)], dtype=np.uint64).view(np.int64)[0])
    gold = int(np.array([GOLDEN], dtype=np.uint64).view(np.int64)[0])
    dtype = dtype or torch.complex64
    out = torch.empty(n, dtype=dtype, device=device)
    m53 = (1 << 53) - 1
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        ctr = torch.arange(2 * s + 1, 2 * e + 1, dtype=torch.int64, device=device)
        z = _t_mix(key + ctr * gold)
        u = ((z >> 11) & m53).to(torch.float64) * (1.0 / 9007199254740992.0)
        r = torch.sqrt(-torch.log1p(-u[0::2]))
        th = 2.0 * np.pi * u[1::2]
        out[s:e] = torch.complex(r * torch.cos(th), r * torch.sin(th)).to(dtype)
        del ctr, z, u, r, th
    return out.reshape(shape)


def config4_torch(Nx: int = 512, Nz: int = 8192, M: int = 96, seed: int = 4, device="cuda",
                  n_sources: int = 4, span: int = 200):
    """config4() generated on a torch device: identical recipe (noise from the counter
    generator, the same seeded sources), returned as a complex64 tensor."""
    import torch
    x = cnormal_torch(seed, 0, (Nx, Nz, M), device)
    u = uniforms(seed, 1, 4 * n_sources)
    m = torch.arange(M, dtype=torch.float64, device=device)
    looks = torch.linspace(-0.5, 0.5, Nx, dtype=torch.float64, device=device)
    for s in range(n_sources):
        th = np.deg2rad(-30.0 + 60.0 * u[4 * s])
        snr = 10.0 + 15.0 * u[4 * s + 1]
        z0 = int(u[4 * s + 2] * (Nz - span))
        amp = np.sqrt(10.0 ** (snr / 10.0))
        sig = torch.from_numpy(cnormal(seed, 10 + s, (span,)) * amp).to(device)
        ph = torch.exp(1j * np.pi * m[None, :] * (np.sin(th) - looks[:, None]))       # [Nx][M]
        x[:, z0:z0 + span, :] += (sig[None, :, None] * ph[:, None, :]).to(torch.complex64)
    return x


def speckle_cube_torch(Nx: int, Nz: int, M: int, seed: int, device="cuda", n_scatterers: int = 10,
                       scatterer_db: float = 30.0, stream: int = 0):
    """speckle_cube() computed with torch on `device` (same recipe, same numbers up to the
    final complex64 rounding)."""
    import torch
    n = cnormal_torch(seed, stream, (Nx, Nz + 4, M), device, dtype=torch.complex128)
    h = AXIAL_PULSE
    cube = torch.zeros((Nx, Nz, M), dtype=torch.complex128, device=device)
    for i in range(5):
        cube += float(h[i]) * n[:, i:i + Nz, :]
    del n
    var = float(np.sum(h * h))
    amp = np.sqrt(10.0 ** (scatterer_db / 10.0) * var)
    for zs, xs, ph in _scatterers(Nx, Nz, seed, n_scatterers, stream):
        a = amp * np.exp(2j * np.pi * ph)
        for i in range(5):
            zz = zs + i - 2
            if 0 <= zz < Nz:
                cube[xs, zz, :] += complex(h[i] * a)
    return cube.to(torch.complex64)


def config3_torch(Nx: int = 256, Nz: int = 4096, seed: int = 3, Mmax: int = 128, device="cuda"):
    """config3() on a torch device: (cube tensor, M_rows, L_rows, off_rows, d_rows numpy)."""
    import torch
    Mz, Lz, off, dz = config3_geometry(Nz, seed, Mmax=Mmax)
    cube = speckle_cube_torch(Nx, Nz, Mmax, seed, device=device)
    m = torch.arange(Mmax, device=device)
    o = torch.from_numpy(off).to(device)[:, None]
    mz = torch.from_numpy(Mz).to(device)[:, None]
    cube *= ((m[None, :] >= o) & (m[None, :] < o + mz))[None].to(cube.dtype)
    return cube, Mz, Lz, off, dz
