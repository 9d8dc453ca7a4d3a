"""Pins the oracle's mixed-radix FFT against numpy.fft (FP64) at every length the
convolutions use, forward and (unnormalised) inverse."""
import numpy as np
import pytest

import oracle_lib as O

LENGTHS = [2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 14, 90, 88, 10]


@pytest.mark.parametrize("n", LENGTHS)
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    f = O.fft(x, -1)
    ref = np.fft.fft(x)
    assert np.abs(f - ref).max() <= 1e-13 * np.abs(ref).max() * np.log2(max(n, 2))
    b = O.fft(x, +1)
    refb = np.fft.ifft(x) * n
    assert np.abs(b - refb).max() <= 1e-13 * np.abs(refb).max() * np.log2(max(n, 2))


def test_fft_brute_dft():
    n = 12
    x = np.arange(n) + 1j * np.arange(n)[::-1]
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    np.testing.assert_allclose(O.fft(x, -1), W @ x, rtol=0, atol=1e-12)
