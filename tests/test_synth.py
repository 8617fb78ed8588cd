"""Synthetic-input generator: the torch implementation reproduces the numpy one exactly,
and the counter stream is sliceable (CPU only)."""
import numpy as np

from cos_amd import synth


def test_torch_generator_matches_numpy_bitwise():
    a = synth.cnormal(7, 3, (5, 7, 11)).astype(np.complex64)
    b = synth.cnormal_torch(7, 3, (5, 7, 11), "cpu", chunk=97).numpy()
    assert np.array_equal(a, b)


def test_config4_torch_matches_numpy():
    x = synth.config4(Nx=4, Nz=300, M=16, seed=4)
    y = synth.config4_torch(Nx=4, Nz=300, M=16, seed=4, device="cpu").numpy()
    assert np.array_equal(x, y)


def test_counter_stream_is_sliceable_and_unit_variance():
    u = synth.uniforms(2, 0, 1000)
    assert np.array_equal(u[300:400], synth.uniforms(2, 0, 100, offset=300))
    n = synth.cnormal(2, 0, (200000,))
    assert abs(np.mean(np.abs(n) ** 2) - 1) < 0.01
    assert abs(np.mean(n.real * n.imag)) < 0.01


def test_speckle_variance_and_scatterers():
    x = synth.speckle_cube(16, 64, 8, seed=9, n_scatterers=0)
    assert abs(np.mean(np.abs(x) ** 2) / 1.625 - 1) < 0.1
    y = synth.speckle_cube(16, 64, 8, seed=9, n_scatterers=3)
    d = np.abs(y - x).max(axis=2)
    assert (d > 10).sum() >= 3                  # scatterer cells stand out, coherent over M


def test_torch_speckle_matches_numpy():
    a = synth.speckle_cube(4, 30, 8, seed=3)
    b = synth.speckle_cube_torch(4, 30, 8, seed=3, device="cpu").numpy()
    assert np.array_equal(a, b)
    c, *_ = synth.config3(Nx=2, Nz=64, seed=3, Mmax=32)
    d, *_ = synth.config3_torch(Nx=2, Nz=64, seed=3, Mmax=32, device="cpu")
    assert np.array_equal(c, d.numpy())
