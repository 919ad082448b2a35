"""CPU models of the potential form's pass sharing: the x^-1 pass's
derivative of two packed real rows, and the byte model bench.py reports for
the one-launch z outputs and the shared y^-1 loads."""
import numpy as np
import pytest

import bench


def _c2r_pair_derivative(xa, xb, L, kst):
    """The specialised x^-1 kernel on one row pair: half spectra xa, xb
    (H = L/2 + 1 values each) packed as v[pos] = za + i zb over the full
    length (Hermitian extension for pos >= H, imaginary parts of the DC and
    Nyquist terms zeroed), then v times i kk with kk = kh(pos) below H (0 at
    the Nyquist index) and -kh(L - pos) from H on, then the unnormalised
    inverse DFT; the two real rows are its real and imaginary parts."""
    H = L // 2 + 1
    v = np.zeros(L, complex)
    for pos in range(L):
        if pos < H:
            za, zb = xa[pos], xb[pos]
            if pos == 0 or 2 * pos == L:
                za, zb = za.real + 0j, zb.real + 0j
        else:
            m = L - pos
            za, zb = np.conj(xa[m]), np.conj(xb[m])
        v[pos] = za + 1j * zb
    kk = np.array([(0.0 if 2 * p == L else kst * p) if p < H else -kst * (L - p) for p in range(L)])
    y = np.fft.ifft(1j * kk * v) * L
    return y.real, y.imag


@pytest.mark.parametrize("L", [8, 64, 256])
def test_packed_row_pair_derivative_matches_each_row(L):
    """Multiplying the packed pair by i kk equals differentiating each row
    on its own (i kh X with the Nyquist term dropped, as the grad operator's
    Hermitian part)."""
    rng = np.random.default_rng(L)
    a, b = rng.standard_normal(L), rng.standard_normal(L)
    dx = 0.37
    kst = 2 * np.pi / (L * dx)
    xa, xb = np.fft.rfft(a), np.fft.rfft(b)
    ga, gb = _c2r_pair_derivative(xa, xb, L, kst)
    kh = kst * np.arange(L // 2 + 1)
    kh[-1] = 0.0                                     # Nyquist: no derivative
    ra = np.fft.irfft(1j * kh * xa, n=L) * L
    rb = np.fft.irfft(1j * kh * xb, n=L) * L
    assert np.allclose(ga, ra, rtol=0, atol=1e-9 * np.abs(ra).max())
    assert np.allclose(gb, rb, rtol=0, atol=1e-9 * np.abs(rb).max())
    # and it is the spectral derivative of a band-limited row
    x = np.arange(L) * dx
    s = np.sin(2 * np.pi * 3 * x / (L * dx))
    gs, _ = _c2r_pair_derivative(np.fft.rfft(s), np.zeros(L // 2 + 1), L, kst)
    assert np.allclose(gs / L, 2 * np.pi * 3 / (L * dx) * np.cos(2 * np.pi * 3 * x / (L * dx)), atol=1e-9)


def _spec(n, bc, eb):
    M = [2 * m if b == 0 else m for m, b in zip(n, bc)]
    H = M[0] // 2 + 1
    return M, H, 2 * eb


@pytest.mark.parametrize("n,bc", [((1024,) * 3, (0, 0, 0)), ((256,) * 3, (1, 1, 0)), ((48, 40, 36), (0, 1, 0))])
def test_potential_byte_model(n, bc):
    """bench.alg_bytes_potential against counts from the grid: one z launch
    reads the spectrum and the kernel octant once and writes each z output;
    shared y^-1 loads read A's z output once per launch; an output taken from
    A's y^-1 output by the x^-1 pass writes no y^-1 spectrum of its own."""
    eb = 4
    M, H, ec = _spec(n, bc, eb)
    outs = ("A", "dx", "dy", "dz", "lap")
    spec_z = H * M[1] * n[2] * ec          # one z-pass spectrum (kept z planes)
    spec_y = H * n[1] * n[2] * ec          # one y^-1 output
    octant = (M[0] // 2 + 1) * (M[1] // 2 + 1) * (M[2] // 2 + 1) * eb
    three = bench.alg_bytes_potential(n, bc, outs, eb)
    one = bench.alg_bytes_potential(n, bc, outs, eb, zlaunches=1)
    # three z transforms (A, dz, lap), each read + octant + write
    assert three["z"] == 3 * (2 * spec_z + octant)
    assert one["z"] == spec_z + octant + 3 * spec_z
    assert three["solve"] - one["solve"] == 2 * (spec_z + octant)
    # y^-1: five outputs, each reading its z spectrum and writing its y spectrum
    assert three["y_inv"] == 5 * (spec_z + spec_y)
    shared = bench.alg_bytes_potential(n, bc, outs, eb, zlaunches=1, ylaunches=3, youtputs=4)
    assert shared["y_inv"] == 3 * spec_z + 4 * spec_y
    # x^-1: d/dx A a second pass over A's rows, which are read once
    assert shared["x_inv"] == one["x_inv"] - spec_y
