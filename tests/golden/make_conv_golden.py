"""Writes tests/golden/conv_golden.npz: numpy-FFT golden vectors for the per-box convolution.

phi: seeded clamped fields on n x n boxes; out = irfft2(G * rfft2(phi)) with the Gaussian
symbol of LSbox::convolutionGeneratorFFTW (src/box.cpp:674-688).  numpy's pocketfft is the
library routine standing in for FFTW, which is not installed here.
"""
import os

import numpy as np

SIZES = (16, 17, 31, 37, 40, 64, 90, 127, 128)
R, PPG, GB = 894, 20, 20


def gaussian_symbol(n, dt, R):
    i = np.arange(n // 2 + 1)
    j = np.arange(n)
    i2 = np.minimum(i, n - i)[None, :]
    j2 = np.minimum(j, n - j)[:, None]
    return np.exp(-np.sqrt(2 / np.pi) * (i2 * i2 + j2 * j2) * 4.0 * dt * (R * R) / (n * n) * np.pi * np.pi) / (n * n)


def main():
    dt = 0.8 / (R * R) * PPG / 2 / 5.0
    delta = GB / R
    rng = np.random.default_rng(1234)
    out = {"dt": dt, "R": R}
    for n in SIZES:
        phi = np.clip(rng.normal(0, delta, (n, n)), -delta, delta)
        spec = np.fft.rfft2(phi) * gaussian_symbol(n, dt, R)
        conv = np.fft.irfft2(spec, s=(n, n)) * (n * n)  # FFTW's c2r is unnormalised
        out[f"phi_{n}"] = phi
        out[f"conv_{n}"] = conv
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "conv_golden.npz"), **out)


if __name__ == "__main__":
    main()
