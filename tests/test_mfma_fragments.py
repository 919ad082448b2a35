"""CPU check of the constants the matrix-core kernels multiply by: the FP16
hi/lo split of the 32-point DFT matrix, per MFMA lane (v_mfma_f32_16x16x32_f16
A operand: lane l holds row l % 16, K = 8 (l / 16) + 0..7).  Pins the layout
against the textbook DFT and the rounding against numpy's binary16."""
import ctypes

import numpy as np

import poissonsolver_amd._capi as C


def fragments():
    lib = C.lib()
    buf = (ctypes.c_ushort * 4096)()
    assert lib.ps_dft32_fragments(buf, 4096) == 0
    assert lib.ps_dft32_fragments(buf, 100) != 0        # too small: error, not a write
    return np.frombuffer(bytes(buf), dtype=np.uint16).view(np.float16).reshape(64, 2, 4, 8)


def test_fragments_are_the_split_dft_matrix():
    f = fragments().astype(np.float64)
    for lane in range(64):
        i, q = lane % 16, lane // 16
        for c in range(2):
            for e8 in range(8):
                m = 4 * q + (e8 >> 1) + 16 * c
                t = 2 * np.pi * i * m / 32
                # Y = F x, F = exp(-2 pi i k m / 32): Yr = cos xr + sin xi, Yi = -sin xr + cos xi
                vr = np.sin(t) if e8 & 1 else np.cos(t)
                vi = np.cos(t) if e8 & 1 else -np.sin(t)
                rh, rl, ih, il = f[lane, c, :, e8]
                assert abs(rh + rl - vr) <= 2.0 ** -22 * max(abs(vr), 2.0 ** -3)
                assert abs(ih + il - vi) <= 2.0 ** -22 * max(abs(vi), 2.0 ** -3)
                # hi is the round-to-nearest binary16 of the value, lo of the rest
                assert rh == np.float64(np.float16(vr)) and ih == np.float64(np.float16(vi))
                assert rl == np.float64(np.float16(vr - rh)) and il == np.float64(np.float16(vi - ih))


def test_split_product_reproduces_a_dft_to_fp32_accuracy():
    # the three-product form A_hi B_hi + A_hi B_lo + A_lo B_hi that the
    # kernels accumulate in FP32, on random data scaled into [-2^10, 2^10]
    f = fragments().astype(np.float64)
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(32) + 1j * rng.standard_normal(32)) * 300.0
    xr = np.empty(64)
    xr[0::2], xr[1::2] = x.real, x.imag                 # K order: (re, im) per point
    xh = xr.astype(np.float16).astype(np.float64)
    xl = (xr - xh).astype(np.float16).astype(np.float64)
    # assemble the 16-row blocks: rows k < 16 from the fragments, K = 16 c + 8 q + e8
    A = np.zeros((2, 2, 16, 64))                          # [re/im][hi/lo][row][K]
    for lane in range(64):
        i, q = lane % 16, lane // 16
        for c in range(2):
            for e8 in range(8):
                K = 32 * c + 8 * q + e8
                A[0, 0, i, K], A[0, 1, i, K], A[1, 0, i, K], A[1, 1, i, K] = f[lane, c, :, e8]
    y = np.fft.fft(x)[:16]
    got_r = A[0, 0] @ xh + A[0, 0] @ xl + A[0, 1] @ xh
    got_i = A[1, 0] @ xh + A[1, 0] @ xl + A[1, 1] @ xh
    err = np.linalg.norm((got_r + 1j * got_i) - y) / np.linalg.norm(y)
    assert err < 1e-6, err


def fragments16():
    lib = C.lib()
    buf = (ctypes.c_ushort * 2048)()
    assert lib.ps_dft16_fragments(buf, 2048) == 0
    assert lib.ps_dft16_fragments(buf, 100) != 0        # too small: error, not a write
    return np.frombuffer(bytes(buf), dtype=np.uint16).view(np.float16).reshape(64, 4, 8)


def test_fragments16_are_the_split_dft16_matrix():
    # the 16-point DFT of the transforms of 512: one K chunk, rows 0..15
    f = fragments16().astype(np.float64)
    for lane in range(64):
        i, q = lane % 16, lane // 16
        for e8 in range(8):
            m = 4 * q + (e8 >> 1)
            t = 2 * np.pi * i * m / 16
            vr = np.sin(t) if e8 & 1 else np.cos(t)
            vi = np.cos(t) if e8 & 1 else -np.sin(t)
            rh, rl, ih, il = f[lane, :, e8]
            assert abs(rh + rl - vr) <= 2.0 ** -22 * max(abs(vr), 2.0 ** -3)
            assert abs(ih + il - vi) <= 2.0 ** -22 * max(abs(vi), 2.0 ** -3)
            assert rh == np.float64(np.float16(vr)) and ih == np.float64(np.float16(vi))
            assert rl == np.float64(np.float16(vr - rh)) and il == np.float64(np.float16(vi - ih))


def test_split_product16_reproduces_a_dft_to_fp32_accuracy():
    f = fragments16().astype(np.float64)
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(16) + 1j * rng.standard_normal(16)) * 300.0
    xr = np.empty(32)
    xr[0::2], xr[1::2] = x.real, x.imag
    xh = xr.astype(np.float16).astype(np.float64)
    xl = (xr - xh).astype(np.float16).astype(np.float64)
    A = np.zeros((2, 2, 16, 32))
    for lane in range(64):
        i, q = lane % 16, lane // 16
        for e8 in range(8):
            K = 8 * q + e8
            A[0, 0, i, K], A[0, 1, i, K], A[1, 0, i, K], A[1, 1, i, K] = f[lane, :, e8]
    y = np.fft.fft(x)
    got_r = A[0, 0] @ xh + A[0, 0] @ xl + A[0, 1] @ xh
    got_i = A[1, 0] @ xh + A[1, 0] @ xl + A[1, 1] @ xh
    err = np.linalg.norm((got_r + 1j * got_i) - y) / np.linalg.norm(y)
    assert err < 1e-6, err
