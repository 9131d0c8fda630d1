"""Writes the FFT golden vectors used by tests/test_oracle_fft.py.

numpy.fft (pocketfft) is an independent FFT implementation: its outputs pin
the oracle's own FFT (which stands in for FFTW 3.3.10, not importable here)
on the reference's conventions -- forward sign -1, unnormalised r2c over
(nz, ny, nx) with x fastest, unnormalised c2r.

    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for n in (8, 16, 12, 64):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"c1d_{n}_in"] = x
        out[f"c1d_{n}_fwd"] = np.fft.fft(x)
        out[f"c1d_{n}_inv"] = np.fft.ifft(x) * n
    for (nx, ny, nz) in ((8, 8, 16), (16, 8, 8), (8, 4, 8)):
        r = rng.standard_normal((nz, ny, nx))
        out[f"r2c_{nx}x{ny}x{nz}_in"] = r
        out[f"r2c_{nx}x{ny}x{nz}_out"] = np.fft.rfftn(r)
        c = np.fft.rfftn(rng.standard_normal((nz, ny, nx)))
        out[f"c2r_{nx}x{ny}x{nz}_in"] = c
        out[f"c2r_{nx}x{ny}x{nz}_out"] = np.fft.irfftn(c, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)
    # lengths with factors 3, 5 and a larger prime (the direct DFT of the
    # oracle; the doubled grids of non-power-of-two shapes)
    for n in (30, 74):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out[f"c1d_{n}_in"] = x
        out[f"c1d_{n}_fwd"] = np.fft.fft(x)
        out[f"c1d_{n}_inv"] = np.fft.ifft(x) * n
    for (nx, ny, nz) in ((12, 10, 6), (10, 6, 14)):
        r = rng.standard_normal((nz, ny, nx))
        out[f"r2c_{nx}x{ny}x{nz}_in"] = r
        out[f"r2c_{nx}x{ny}x{nz}_out"] = np.fft.rfftn(r)
        c = np.fft.rfftn(rng.standard_normal((nz, ny, nx)))
        out[f"c2r_{nx}x{ny}x{nz}_in"] = c
        out[f"c2r_{nx}x{ny}x{nz}_out"] = np.fft.irfftn(c, s=(nz, ny, nx), axes=(0, 1, 2)) * (nx * ny * nz)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
