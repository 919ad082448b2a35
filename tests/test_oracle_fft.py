"""Pins the oracle's restatement of the reference's Glassman FFT
(pencil/fft.cpp, pencil/ifft.cpp): forward sign, unnormalised inverse, output
order, every factor path (prime, composite, repeated factors)."""
import numpy as np
import pytest

from oracle import oracle as O


def dft(x, sign):
    n = x.shape[-1]
    k = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ w.T


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 20, 25, 27, 30, 32, 49, 64, 96, 100])
def test_glassman_matches_textbook_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    f = O.glassman_fft(x)
    fi = O.glassman_fft(x, inverse=True)
    assert np.allclose(f, dft(x, -1), rtol=0, atol=1e-12 * max(1, n))
    assert np.allclose(fi, dft(x, +1), rtol=0, atol=1e-12 * max(1, n))
    # the inverse is unnormalised: ifft(fft(x)) = n x
    assert np.allclose(O.glassman_fft(f, inverse=True), n * x, atol=1e-11 * n)


def test_fft_shift_swaps_halves():
    x = np.arange(8.0)
    assert list(O.fft_shift(x)) == [4, 5, 6, 7, 0, 1, 2, 3]


def test_pipeline_with_glassman_equals_numpy_fft():
    rng = np.random.default_rng(5)
    n = (6, 5, 4)
    b = [rng.standard_normal((n[2], n[1], n[0])) for _ in range(3)]
    a1, r1 = O.solve(n, (0, 0, 0), (0.1, 0.12, 0.09), 4, b, op="curl", reproject=True)
    a2, r2 = O.solve(n, (0, 0, 0), (0.1, 0.12, 0.09), 4, b, op="curl", reproject=True,
                     use_glassman=True)
    assert O.rel_l2(a2, a1) < 1e-13
    assert O.rel_l2(r2, r1) < 1e-13
