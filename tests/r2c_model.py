"""numpy model of the product's 5-pass pipeline (test helper, CPU only).

It restates what the HIP kernels compute, so its agreement with the oracle
(a C2C restatement of the reference) checks the design's mathematics on the
CPU before any kernel runs:

  * x is transformed real-to-complex (kx in [0, Mx/2]);
  * Ghat * zeta / (Mx My Mz) is folded into one real array that is even in
    each wavenumber, so only the octant is stored;
  * the reference takes the real part of a complex inverse; in the half
    spectrum that is the Hermitian part of every operator coefficient.  For a
    single derivative i*k_d the Hermitian part is i*k_d with k_d set to 0 on
    the Nyquist index of d; for the product k_d*k_e of the reprojection it is
    kh_d*kh_e + kn_d*kn_e (kh: Nyquist zeroed, kn: Nyquist only);
  * the inverse x transform keeps Re() of the DC and Nyquist bins.
"""
import numpy as np

from oracle import oracle as O


def kappas(M, dx):
    m = np.arange(M)
    mm = np.where(2 * m < M, m, m - M).astype(np.float64)
    k = 2.0 * np.pi * mm / (M * dx)
    nyq = (2 * m == M)
    kh = np.where(nyq, 0.0, k)
    kn = np.where(nyq, k, 0.0)
    return k, kh, kn


def geff_octant(greens, reg):
    g = greens.g3d.real.copy()
    if greens.zeta is not None:
        g = g * greens.zeta
    Mx, My, Mz = greens.M
    g /= float(Mx * My * Mz)
    return g[: Mz // 2 + 1, : My // 2 + 1, : Mx // 2 + 1]


def expand_octant(oct_, M):
    Mx, My, Mz = M
    iz = np.minimum(np.arange(Mz), Mz - np.arange(Mz))
    iy = np.minimum(np.arange(My), My - np.arange(My))
    ix = np.arange(Mx // 2 + 1)
    return oct_[np.ix_(iz, iy, ix)]


def model_solve(ncell, bc, dx, reg, rhs, op="plain", reproject=False, want_rhs=False):
    greens = O.greens3d(ncell, bc, dx, reg)
    M = greens.M
    Mx, My, Mz = M
    nx, ny, nz = ncell
    G = expand_octant(geff_octant(greens, reg), M)
    kx, khx, knx = kappas(Mx, dx[0])
    ky, khy, kny = kappas(My, dx[1])
    kz, khz, knz = kappas(Mz, dx[2])
    H = Mx // 2 + 1
    KH = [khx[None, None, :H], khy[None, :, None], khz[:, None, None]]
    KN = [knx[None, None, :H], kny[None, :, None], knz[:, None, None]]
    KF = [kx[None, None, :H], ky[None, :, None], kz[:, None, None]]

    def fwd(b):
        a = np.zeros((nz, ny, Mx))
        a[:, :, :nx] = b
        s = np.fft.rfft(a, axis=2)
        t = np.zeros((nz, My, H), complex)
        t[:, :ny] = s
        t = np.fft.fft(t, axis=1)
        u = np.zeros((Mz, My, H), complex)
        u[:nz] = t
        return np.fft.fft(u, axis=0)

    def inv(f):
        f = (np.fft.ifft(f, axis=0) * Mz)[:nz]
        f = (np.fft.ifft(f, axis=1) * My)[:, :ny]
        return (np.fft.irfft(f, n=Mx, axis=2) * Mx)[:, :, :nx]

    S = [None if b is None else fwd(np.asarray(b, float)) for b in rhs]
    i = 1j
    if op == "plain":
        L = [None if s is None else G * s for s in S]
    elif op == "grad":
        L = [i * KH[d] * G * S[d] for d in range(3)]
    elif op == "div":
        L = [G * i * (KH[0] * S[0] + KH[1] * S[1] + KH[2] * S[2]), None, None]
    elif op == "curl":
        L = [G * i * (KH[1] * S[2] - KH[2] * S[1]),
             G * i * (KH[2] * S[0] - KH[0] * S[2]),
             G * i * (KH[0] * S[1] - KH[1] * S[0])]
    lhs = [None if l is None else inv(l) for l in L]
    if not want_rhs:
        return lhs, None
    # rhs filter: zeta/N * S, then - Geff * (kh_d kh.S + kn_d kn.S)
    N = float(Mx * My * Mz)
    if greens.zeta is not None:
        Z = greens.zeta[:, :, :H] / N
    else:
        Z = np.full(G.shape, 1.0 / N)
    R = [None if s is None else Z * s for s in S]
    if reproject and all(s is not None for s in S):
        dh = KH[0] * S[0] + KH[1] * S[1] + KH[2] * S[2]
        dn = KN[0] * S[0] + KN[1] * S[1] + KN[2] * S[2]
        R = [R[d] - G * (KH[d] * dh + KN[d] * dn) for d in range(3)]
    rb = [None if r is None else inv(r) for r in R]
    return lhs, rb
