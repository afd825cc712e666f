"""Pins of the oracle's FFT (radix-2 Stockham) and pairwise sums (CPU only).

The reference's FFTs are MKL/FFTW calls (closed library); their rounding is
not reproducible, so the oracle's own FFT is pinned to the DFT definition
and numpy.fft, and the GPU kernels are pinned bit-for-bit to the oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", [4, 8, 16, 64, 256, 512, 1024, 4096])
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r, i = O.fft(x.real, x.imag, -1)
    y = np.fft.fft(x)
    assert np.abs((r + 1j * i) - y).max() <= 8e-16 * math.log2(n) * np.abs(y).max() + 1e-300
    r, i = O.fft(x.real, x.imag, +1)
    y = np.fft.ifft(x) * n
    assert np.abs((r + 1j * i) - y).max() <= 8e-16 * math.log2(n) * np.abs(y).max()


def test_fft_dft_definition_long_double():
    n = 16
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * (np.outer(k, k) % n).astype(np.longdouble) / n)
    y = (W * x[None, :]).sum(axis=1)
    r, i = O.fft(x.real, x.imag, -1)
    assert np.abs((r + 1j * i) - y.astype(np.complex128)).max() < 1e-14


def test_fft_roundtrip_and_real_input_hermitian():
    n = 64
    rng = np.random.default_rng(2)
    x = rng.standard_normal(n)
    r, i = O.fft(x, np.zeros(n), -1)
    # Hermitian symmetry of a real input's spectrum (exact up to rounding)
    assert np.allclose(r[1:], r[1:][::-1], atol=1e-13) and np.allclose(i[1:], -i[1:][::-1], atol=1e-13)
    r2, i2 = O.fft(r, i, +1)
    assert np.allclose(r2 / n, x, atol=1e-14) and np.abs(i2).max() < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 100, 513])
def test_tree_sum(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    s = O.tree_sum(x)
    assert abs(s - math.fsum(x)) <= 4e-16 * math.log2(max(n, 2)) * np.abs(x).sum()
    xi = rng.integers(-1000, 1000, n).astype(np.float64)
    assert O.tree_sum(xi) == xi.sum()  # exact for small integers


def test_tree_sum_pairing():
    # pairwise order: ((x0+x2)+(x1+x3)) for n = 4
    x = np.array([1.0, 1e16, -1.0, -1e16])
    assert O.tree_sum(x) == (x[0] + x[2]) + (x[1] + x[3])
