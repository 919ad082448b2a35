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


def table16(kind):
    lib = C.lib()
    buf = (ctypes.c_ushort * 2048)()
    assert lib.ps_dft16_table(buf, 2048, kind) == 0
    assert lib.ps_dft16_table(buf, 100, kind) != 0       # too small: error, not a write
    assert lib.ps_dft16_table(buf, 2048, 3) != 0         # no such table
    return np.frombuffer(bytes(buf), dtype=np.uint16).view(np.float16).reshape(64, 4, 8).astype(np.float64)


def table16_exponent(kind, i, k):
    """W32 exponent of entry (i, k) of table `kind`: 0 the 16-point DFT, 1 P, 2 Q."""
    return [2 * i * k, k * (2 * i + 1), i * (2 * k + 1)][kind]


def test_tables16_are_split_w32_powers():
    # kind 0 equals the DFT-16 fragments; P and Q are W32^(k(2i+1)), W32^(i(2k+1))
    assert np.array_equal(table16(0), fragments16().astype(np.float64))
    for kind in (1, 2):
        f = table16(kind)
        for lane in range(64):
            i, q = lane % 16, lane // 16
            for e8 in range(8):
                k = 4 * q + (e8 >> 1)
                t = 2 * np.pi * table16_exponent(kind, i, k) / 32
                vr = np.sin(t) if e8 & 1 else np.cos(t)
                vi = np.cos(t) if e8 & 1 else -np.sin(t)
                rh, rl, ih, il = f[lane, :, e8]
                assert rh == np.float64(np.float16(vr)) and ih == np.float64(np.float16(vi))
                assert rl == np.float64(np.float16(vr - rh)) and il == np.float64(np.float16(vi - ih))


def table16_matrix(kind):
    """The complex 16 x 16 matrix a table holds (hi + lo), rebuilt from its
    (re, im) K slots: entry (i, k) = vr + i vi of the real-part slot."""
    f = table16(kind)
    M = np.zeros((16, 16), complex)
    for lane in range(64):
        i, q = lane % 16, lane // 16
        for e in range(4):
            k = 4 * q + e
            M[i, k] = (f[lane, 0, 2 * e] + f[lane, 1, 2 * e]) + 1j * (f[lane, 2, 2 * e] + f[lane, 3, 2 * e])
    return M


def register_pencil_pass(x, g, E, P, Q):
    """The z pass of one pencil as the register-pencil matrix-core kernel
    factors it: 512 input points, the transform of 1024 as 32 x 32 with every
    32-point stage split into even/odd 16-point products (E = W16^(ik),
    P = W32^(k(2i+1)), Q = W32^(i(2k+1))), the k-space multiplier g[kz] between
    the forward and the inverse transform, the first 512 outputs kept."""
    W = np.exp(-2j * np.pi / 1024)
    x2 = x.reshape(16, 32).T                                   # x2[b, m] = x[b + 32 m]
    Y = np.zeros((32, 32), complex)                            # Y[b, k1]
    Y[:, 0::2] = x2 @ E.T                                      # even k1 = 2 k': sum_m W16^(k' m)
    Y[:, 1::2] = x2 @ P.T                                      # odd: sum_m W32^(m (2k'+1))
    b = np.arange(32)[:, None]
    k1 = np.arange(32)[None, :]
    Y = Y * W ** (b * k1)
    Ev = E @ Y[0::2, :]                                        # [k2, k1] from even b
    Od = Q @ Y[1::2, :]                                        # odd b: W32^(k2 (2 b' + 1))
    X = np.vstack([Ev + Od, Ev - Od])                          # rows k2 = 0..31
    kz = k1 + 32 * np.arange(32)[:, None]
    Z = X * g[kz]
    A = np.conj(Z[:16] + Z[16:])                               # [k2 < 16, k1]
    B = np.conj(Z[:16] - Z[16:])
    V = np.zeros((32, 32), complex)                            # V[k1, n2]
    V[:, 0::2] = (E @ A).T
    V[:, 1::2] = (P @ B).T
    n2 = k1
    T = V * (W ** (b * n2))                                    # b plays k1 here
    out = np.conj(E @ T[0::2, :] + Q @ T[1::2, :])             # [n1 < 16, n2]
    return out.reshape(-1)                                     # index 32 n1 + n2


def test_register_pencil_factorisation_matches_fft_convolution():
    # exact W32 powers: the factorisation itself (1e-12); the FP16 hi + lo
    # tables: their rounding (< 1e-6, the mixed form's tolerance)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(512) + 1j * rng.standard_normal(512)
    kz = np.arange(1024)
    g = np.exp(-0.5 * (np.minimum(kz, 1024 - kz) / 90.0) ** 2)
    ref = np.fft.ifft(np.fft.fft(x, 1024) * g)[:512] * 1024
    i = np.arange(16)[:, None]
    k = np.arange(16)[None, :]
    w32 = np.exp(-2j * np.pi / 32)
    exact = [w32 ** table16_exponent(kind, i, k) for kind in range(3)]
    got = register_pencil_pass(x, g, *exact)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    split = [table16_matrix(kind) for kind in range(3)]
    err = np.linalg.norm(register_pencil_pass(x, g, *split) - ref) / np.linalg.norm(ref)
    assert err < 1e-6, err
    # a plausible slip (P where Q belongs) is caught
    bad = register_pencil_pass(x, g, exact[0], exact[2], exact[1])
    assert np.linalg.norm(bad - ref) / np.linalg.norm(ref) > 1e-2
