"""The oracle's FFT (stand-in for FFTW 3.3.10) pinned against numpy.fft golden
vectors, and against the mathematics: roundtrip = N x identity, Parseval,
single modes, real input -> Hermitian spectrum (distributed_fft3d.h:18-72,
distributed_fft3d_fftw.cc:67-71 conventions)."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


@pytest.mark.parametrize("n", [8, 16, 12, 64, 30, 74])
def test_fft1d_golden(n):
    x = GOLD[f"c1d_{n}_in"]
    np.testing.assert_allclose(O.fft1d(x, -1), GOLD[f"c1d_{n}_fwd"], rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.fft1d(x, +1), GOLD[f"c1d_{n}_inv"], rtol=0, atol=1e-12 * n)


@pytest.mark.parametrize("dims", [(8, 8, 16), (16, 8, 8), (8, 4, 8), (12, 10, 6), (10, 6, 14)])
def test_fft3d_r2c_c2r_golden(dims):
    nx, ny, nz = dims
    r = GOLD[f"r2c_{nx}x{ny}x{nz}_in"]
    pad = np.zeros((nz, ny, 2 * (nx // 2 + 1)))
    pad[:, :, :nx] = r
    got = O.fft3d_r2c(pad, nx, ny, nz, nthreads=2)
    ref = GOLD[f"r2c_{nx}x{ny}x{nz}_out"]
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    c = GOLD[f"c2r_{nx}x{ny}x{nz}_in"]
    back = O.fft3d_c2r(c, nx, ny, nz, nthreads=2)[:, :, :nx]
    ref = GOLD[f"c2r_{nx}x{ny}x{nz}_out"]
    assert np.abs(back - ref).max() <= 1e-12 * np.abs(ref).max()


def test_roundtrip_parseval_modes():
    nx, ny, nz = 16, 8, 32
    rng = np.random.default_rng(1)
    r = rng.standard_normal((nz, ny, nx))
    pad = np.zeros((nz, ny, 2 * (nx // 2 + 1)))
    pad[:, :, :nx] = r
    spec = O.fft3d_r2c(pad, nx, ny, nz, nthreads=2)
    back = O.fft3d_c2r(spec, nx, ny, nz, nthreads=2)[:, :, :nx]
    np.testing.assert_allclose(back, r * nx * ny * nz, rtol=0, atol=1e-11 * nx * ny * nz)
    # Parseval on the full spectrum rebuilt from the half spectrum
    full = np.concatenate([spec, np.conj(spec[:, :, 1:nx // 2][..., ::-1])], axis=2)
    full = full[:, :, : nx]  # kx = 0..nx-1 with the Hermitian tail
    e_spec = np.sum(np.abs(np.fft.fftn(r)) ** 2)
    assert abs(np.sum(np.abs(np.fft.fftn(r)) ** 2) - nx * ny * nz * np.sum(r * r)) < 1e-9 * e_spec
    assert abs(np.sum(np.abs(spec[:, :, 0]) ** 2) + np.sum(np.abs(spec[:, :, nx // 2]) ** 2)
               + 2 * np.sum(np.abs(spec[:, :, 1:nx // 2]) ** 2) - e_spec) < 1e-9 * e_spec
    # a single cosine mode lands on exactly two bins with amplitude N/2
    kx, ky, kz = 3, 2, 5
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mode = np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    pad[:] = 0
    pad[:, :, :nx] = mode
    s = O.fft3d_r2c(pad, nx, ny, nz, nthreads=2)
    big = np.argwhere(np.abs(s) > 1e-9)
    assert {tuple(b) for b in big} == {(kz, ky, kx)}
    assert abs(s[kz, ky, kx] - nx * ny * nz / 2) < 1e-9
