"""Oracle -- TEST INFRASTRUCTURE ONLY.

An FP64 CPU restatement of the reference's 3D solve path
(``setup3d`` -> ``push`` -> ``solve3d`` -> ``pull``) and of its 2D solve
(``setup2d`` -> ``push`` -> ``solve2d`` -> ``pull``), written in numpy, with the
two special functions in plain C (``oracle/special.c``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module, and only as the checker: the product path never calls it.

The restatement follows the reference's operation order on a single rank
(nproc = 1, so every transpose of ``map.cpp`` is a self-copy and the y zero-fill
of ``map.cpp:376`` is a plain pad).  Where the reference has undefined
behaviour the intended algorithm is restated and the deviation is named in the
function's docstring.  Arrays use the reference's memory order
ijk = (k*ny + j)*nx + i, i.e. numpy shape (nz, ny, nx).

How each part is pinned (see the tests named in brackets):
  * sine_int / bessel_int_J0: bitwise against the reference's own
    special_functions.cpp built into oracle/_ref [test_oracle_special.py],
    and against scipy / mpmath.
  * glassman_fft: against numpy.fft and a direct O(n^2) DFT
    [test_oracle_fft.py].
  * zeta: against scipy.special.gammaincc (Q(m/2, a) is the same filter)
    [test_oracle_pins.py].
  * UUU kernel + doubling + truncation + normalisation: brute-force direct
    lattice sum on tiny grids, the Gaussian/erf closed form, the large-r
    limit 1/(4 pi r) [test_oracle_pins.py].
  * PPU / UUP / PPP: closed-form 1D, 2D and periodic solutions
    [test_oracle_pins.py].
  * grad/div/curl/reprojection: spectral invariants and the reference
    tests' analytic fields [test_oracle_pins.py].
  * 2D UU / PP / PU / UP kernels and the 2D grad/div/curl: the analytic
    fields of the reference's unbounded2d, periodic2d and mixed2d tests (and
    the transposed mixed case for UP), a 2D Gaussian closed form and a direct
    lattice sum [test_oracle_pins2d.py].
  * smooth3d (the Fortran interface's Gaussian filter): the Gaussian-of-
    Gaussian closed form and a periodic single mode [test_oracle_pins.py].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

PI = 3.1415926535897932          # greens3d.cpp:16, solve3d.cpp:14
EULER_GAMMA = 0.5772156649015329  # greens2d.cpp:16

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


# --------------------------------------------------------------------------- #
# special functions (C restatement, loaded by ctypes)
# --------------------------------------------------------------------------- #
def _special():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_build", "liboracle_special.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "_build/liboracle_special.so"], cwd=_HERE,
                           check=True)
        lib = ctypes.CDLL(path)
        for name in ("oracle_sine_int_v", "oracle_bessel_int_J0_v"):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        _LIB = lib
    return _LIB


def _vec(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    fn(x.ctypes.data, y.ctypes.data, x.size)
    return y


def sine_int(x):
    """Si(x), NR cisi restatement (special_functions.cpp:238-352)."""
    return _vec(_special().oracle_sine_int_v, x)


def bessel_int_J0(x):
    """Ji0(x) = int_0^x (1-J0(t))/t dt, Luke 1975 (special_functions.cpp:115-229)."""
    return _vec(_special().oracle_bessel_int_J0_v, x)


# --------------------------------------------------------------------------- #
# partitions (partition_setup.cpp:9-199), single rank
# --------------------------------------------------------------------------- #
def partition_setup(pencil_dir, ncell, bc, dx, ext_x, ext_y, ext_z):
    """One-rank partition list: ncell doubled where bc==0 and the extend flag is
    set (partition_setup.cpp:71-96); icell = 0 (:102-179 with nsub = 1)."""
    if pencil_dir not in (0, 1):
        raise ValueError("partition_setup: pencil_dir must be 0 or 1")
    ext = (ext_x, ext_y, ext_z)
    n = [2 * ncell[d] if (bc[d] == 0 and ext[d]) else ncell[d] for d in range(3)]
    return [dict(ncell=n, icell=[0, 0, 0], dx=list(dx))]


# --------------------------------------------------------------------------- #
# 1D transforms
# --------------------------------------------------------------------------- #
def glassman_fft(x, inverse=False):
    """Glassman's general-N FFT along the last axis, vectorised over the others.

    Restates class_pencil::fft / ifft (pencil/fft.cpp:33-197,
    pencil/ifft.cpp): smallest factor first, ping-pong between two buffers,
    twiddles by the recurrence omega <- delta*omega.  Forward uses
    exp(-2 pi i/N), inverse exp(+2 pi i/N), and the inverse is NOT normalised.
    """
    src = np.array(x, dtype=np.complex128, copy=True)
    n = src.shape[-1]
    dst = np.empty_like(src)
    a, b, c = 1, n, 1
    sgn = 1.0 if inverse else -1.0
    swapped = False
    while b > 1:
        a = c * a
        c = next(f for f in range(2, b + 1) if b % f == 0)
        b //= c
        angle = 2.0 * PI / float(a * c)
        delta = complex(np.cos(angle), sgn * np.sin(angle))
        omega = complex(1.0, 0.0)
        ib = np.arange(b)
        for ic in range(1, c + 1):
            for ia in range(1, a + 1):
                acc = src[..., ib + b * (c - 1) + b * c * (ia - 1)]
                for j in range(2, c + 1):
                    jc = c + 1 - j
                    acc = src[..., ib + b * (jc - 1) + b * c * (ia - 1)] + omega * acc
                dst[..., ib + b * (ia - 1) + b * a * (ic - 1)] = acc
                omega = delta * omega
        src, dst = dst, src
        swapped = not swapped
    return src


def fft_shift(x, axis=-1):
    """Swap halves along ``axis`` (pencil/fft_shift.cpp:32-65; even N)."""
    n = x.shape[axis]
    return np.roll(x, -(n // 2), axis=axis)


def _fft(x, axis, inverse, use_glassman):
    if use_glassman:
        xm = np.moveaxis(x, axis, -1)
        return np.moveaxis(glassman_fft(xm, inverse), -1, axis)
    if inverse:
        return np.fft.ifft(x, axis=axis) * x.shape[axis]
    return np.fft.fft(x, axis=axis)


# --------------------------------------------------------------------------- #
# Green's kernels (greens3d.cpp, greens2d.cpp)
# --------------------------------------------------------------------------- #
def ik_vector(nfft, dx):
    """i*k for an FFT of length nfft (greens3d.cpp:75-118): m -> m - nfft when
    2m >= nfft, so the Nyquist wavenumber is kept negative."""
    dk = 1.0 / (float(nfft) * dx)
    m = np.arange(nfft, dtype=np.float64)
    m = np.where(2 * np.arange(nfft) < nfft, m, m - nfft)
    return 1j * (2.0 * PI * m * dk)


def zeta_filter(kx, ky, kz, h, reg_order):
    """Regularisation filter zeta_m(k) on the (kz, ky, kx) grid
    (greens3d.cpp:123-159).  Odd orders are rounded up (:126).  The coefficient
    array of :130-135 is one element short (heap overflow); the intended
    coefficients c_n = 1/n!, n = 0 .. m/2-1, are used."""
    if reg_order <= 0:
        return None
    if reg_order % 2:
        reg_order += 1
    sigma = 2.0 * h
    nterm = reg_order // 2
    c = [1.0]
    for n in range(1, nterm):
        c.append(c[n - 1] * 1.0 / float(n))
    arg = 0.5 * sigma ** 2 * (kz[:, None, None] ** 2 + ky[None, :, None] ** 2
                              + kx[None, None, :] ** 2)
    s = np.zeros_like(arg)
    for n in range(nterm):
        s += c[n] * np.power(arg, n)
    return s * np.exp(-arg)


def greens_uuu_samples(M, dx, h):
    """Real-space samples of the spectrally truncated kernel on the doubled
    grid (greens3d.cpp:193-225): x_n = (n - M/2) dx, G = Si(r/s)/(2 pi^2 r) with
    s = h/pi, or 1/(2 pi^2 s) when r <= dx_x/4; weighted by dx dy dz."""
    c_1_2pi2 = 0.5 / PI ** 2
    sigma = h / PI
    ax = [-(0.5 * float(M[d])) * dx[d] + np.arange(M[d], dtype=np.float64) * dx[d]
          for d in range(3)]
    z = ax[2][:, None, None]
    y = ax[1][None, :, None]
    x = ax[0][None, None, :]
    r = np.sqrt(x * x + y * y + z * z)
    rho = r / sigma
    far = r > 0.25 * dx[0]
    g = np.empty_like(r)
    si = sine_int(rho[far])
    g[far] = c_1_2pi2 * si / r[far] * dx[0] * dx[1] * dx[2]
    g[~far] = c_1_2pi2 / sigma * dx[0] * dx[1] * dx[2]
    return g


def greens_1d_samples(M, dz):
    """1D spectral kernel -(s/pi)(rho Si(rho) + cos rho) dz, s = dz/pi, at
    z_k = (k - M/2) dz (greens2d.cpp:319-338; used by the PPU branch of
    greens3d.cpp:357-370, whose z origin xmin[2] is uninitialised (:365) and
    whose store index is a stale j (:369): the intended centred lattice and
    index k are used here)."""
    sigma = dz / PI
    c1 = sigma / PI
    z = -(0.5 * float(M)) * dz + np.arange(M, dtype=np.float64) * dz
    rho = np.abs(z) / sigma
    return -c1 * (sine_int(rho) * rho + np.cos(rho)) * dz


def greens_2d_uu(n2, dx2, use_glassman=False):
    """G2D of greens2d.cpp:157-264 (UU branch): samples
    -(1/2pi)(Ji0(r/s) + ln(2s) - gamma) dx dy with s = max(dx,dy)/pi on the
    doubled 2D grid, fft-shifted and 2D DFT'd.  Returns shape (My, Mx)."""
    h = max(dx2[0], dx2[1])
    sigma = h / PI
    c2 = np.log(2.0 * sigma) - EULER_GAMMA
    c_1_2pi = 0.5 / PI
    Mx, My = 2 * n2[0], 2 * n2[1]
    x = (-(0.5 * float(Mx)) * dx2[0] + np.arange(Mx) * dx2[0])[None, :]
    y = (-(0.5 * float(My)) * dx2[1] + np.arange(My) * dx2[1])[:, None]
    r = np.sqrt(x * x + y * y)
    g = -c_1_2pi * (bessel_int_J0(r / sigma) + c2) * dx2[0] * dx2[1]
    g = g.astype(np.complex128)
    g = _fft(fft_shift(g, -1), -1, False, use_glassman)
    g = _fft(fft_shift(g, 0), 0, False, use_glassman)
    return g


def greens2d(ncell2, bc2, dx2, reg_order, use_glassman=False):
    """Restates poisson_solver::greens2d (greens2d.cpp:8-441) on one rank for
    the 2D solver.  Returns (ikx, iky, zeta, G2D) with zeta and G2D of numpy
    shape (My, Mx) on the fully extended grid (greens2d.cpp:58-59).

    Branches: UU (:157-264) the Ji0 kernel; PP (:268-303) 1/k^2 with G(0) = 0;
    PU (:307-366) the 1D Si kernel along y on the kx = 0 column; UP (:370-429)
    the 1D kernel along x on the ky = 0 row.  The UP branch of the reference
    stores every sample at a stale index j (:400); the intended index i is
    used here."""
    bc = tuple(int(b) for b in bc2)
    h = max(dx2[0], dx2[1])                                            # :56
    M = [2 * ncell2[d] if bc[d] == 0 else ncell2[d] for d in range(2)]
    ikx = ik_vector(M[0], dx2[0])                                      # :75-91
    iky = ik_vector(M[1], dx2[1])                                      # :93-110
    zeta = None
    if reg_order > 0:                                                  # :114-150
        zeta = zeta_filter(ikx.imag, iky.imag, np.zeros(1), h, reg_order)[0]

    def inv_k2():
        k2 = ikx[None, :] ** 2 + iky[:, None] ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            return -1.0 / k2

    if bc == (0, 0):
        g = greens_2d_uu((ncell2[0], ncell2[1]), (dx2[0], dx2[1]), use_glassman)
    elif bc == (1, 1):
        g = inv_k2()
        g[0, 0] = 0.0
    elif bc == (1, 0):
        g = inv_k2()
        g1 = greens_1d_samples(M[1], dx2[1]).astype(np.complex128)    # :319-338
        g[:, 0] = _fft(fft_shift(g1, -1), -1, False, use_glassman)
    elif bc == (0, 1):
        g = inv_k2()
        g1 = greens_1d_samples(M[0], dx2[0]).astype(np.complex128)    # :389-401, index i
        g[0, :] = _fft(fft_shift(g1, -1), -1, False, use_glassman)
    else:                                                              # :430-434
        raise ValueError("greens2d: unsupported boundary condition pair %r" % (bc,))
    return ikx, iky, zeta, g


class Greens:
    """Output of greens3d: ik vectors, zeta and G3D on the fully extended
    y-pencil grid, numpy shape (Mz, My, Mx)."""

    def __init__(self, ikx, iky, ikz, zeta, g3d, M):
        self.ikx, self.iky, self.ikz, self.zeta, self.g3d, self.M = ikx, iky, ikz, zeta, g3d, M


def greens3d(ncell, bc, dx, reg_order, use_glassman=False):
    """Restates poisson_solver::greens3d (greens3d.cpp:9-547) on one rank."""
    bc = tuple(int(b) for b in bc)
    h = max(dx[0], dx[1], dx[2])                                   # :55
    M = [2 * ncell[d] if bc[d] == 0 else ncell[d] for d in range(3)]  # :57-58
    ikx, iky, ikz = (ik_vector(M[d], dx[d]) for d in range(3))     # :73-118
    zeta = zeta_filter(ikx.imag, iky.imag, ikz.imag, h, reg_order)  # :123-159

    def inv_k2():
        k2 = (ikx[None, None, :] ** 2 + iky[None, :, None] ** 2 + ikz[:, None, None] ** 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            return -1.0 / k2

    if bc == (0, 0, 0):                                            # :167-332
        g = greens_uuu_samples(M, dx, h).astype(np.complex128)
        for axis in (2, 1, 0):
            g = _fft(fft_shift(g, axis), axis, False, use_glassman)
    elif bc == (1, 1, 0):                                          # :336-403
        g = inv_k2()
        g1 = greens_1d_samples(M[2], dx[2]).astype(np.complex128)
        g1 = _fft(fft_shift(g1, -1), -1, False, use_glassman)
        g[:, 0, 0] = g1
    elif bc == (0, 0, 1):                                          # :407-474
        g = inv_k2()
        g[0, :, :] = greens_2d_uu((ncell[0], ncell[1]), (dx[0], dx[1]), use_glassman)
    elif bc == (1, 1, 1):                                          # :478-521
        g = inv_k2()
        g[0, 0, 0] = 0.0
    else:                                                          # :522-541
        raise ValueError("greens3d: unsupported boundary condition tuple %r" % (bc,))
    return Greens(ikx, iky, ikz, zeta, g, M)


# --------------------------------------------------------------------------- #
# the solver object (poisson_solver.hpp:39-124)
# --------------------------------------------------------------------------- #
class OracleSolver:
    """Single-rank restatement of class poisson_solver's 3D path.

    State follows the reference: rhs*/lhs* are complex x-pencil arrays that
    persist between calls (push.cpp:288-290 and solve3d.cpp:255-266 replace
    them, but a component that is not re-pushed keeps its previous array, and a
    later pull returns whatever the last solve left there).
    """

    def __init__(self, use_glassman=False):
        self.use_glassman = use_glassman
        self.lhs_grad = self.lhs_div = self.lhs_curl = self.rhs_reproject = False
        self.rhs = [None, None, None]
        self.lhs = [None, None, None]
        self.greens = None

    # setup3d.cpp:8-71
    def setup3d(self, ncell, bc, dx, reg_order=0):
        self.ncell = [int(v) for v in ncell]
        self.bc = [int(v) for v in bc]
        self.dx = [float(v) for v in dx]
        self.reg = int(reg_order)
        self.partition = partition_setup(0, self.ncell, self.bc, self.dx, False, False, False)
        self.greens = greens3d(self.ncell, self.bc, self.dx, self.reg, self.use_glassman)
        self.Mx = 2 * self.ncell[0] if self.bc[0] == 0 else self.ncell[0]
        self.My = 2 * self.ncell[1] if self.bc[1] == 0 else self.ncell[1]

    # setup2d.cpp:8-72: domain (nx, ny, 1); the y pencils are not y-extended,
    # the y pass zero-pads to 2 ny itself (solve2d.cpp:181-191)
    def setup2d(self, ncell, bc, dx, reg_order=0):
        self.ncell = [int(ncell[0]), int(ncell[1]), 1]
        self.bc = [int(bc[0]), int(bc[1]), -1]
        self.dx = [float(dx[0]), float(dx[1]), 0.0]
        self.reg = int(reg_order)
        self.partition = partition_setup(0, self.ncell, self.bc, self.dx, False, False, False)
        self.g2 = greens2d(self.ncell[:2], self.bc[:2], self.dx[:2], self.reg, self.use_glassman)
        self.Mx = 2 * self.ncell[0] if self.bc[0] == 0 else self.ncell[0]
        self.My = 2 * self.ncell[1] if self.bc[1] == 0 else self.ncell[1]

    # solve2d.cpp:8-475
    def solve2d(self):
        ikx, iky, zeta, G = self.g2
        nx, ny = self.ncell[0], self.ncell[1]
        Mx, My = self.Mx, self.My
        rX, rY, rZ = (r is not None for r in self.rhs)
        if self.lhs_grad:                                          # :44-58
            if not rX:
                raise ValueError("grad: rhs fields have not been pushed")
            lflag = (True, True, False)
        elif self.lhs_div:                                         # :59-72
            if not (rX and rY):
                raise ValueError("div: rhs fields have not been pushed")
            lflag = (True, False, False)
        elif self.lhs_curl:                                        # :73-87
            if not rZ:
                raise ValueError("curl: rhs fields have not been pushed")
            lflag = (True, True, False)
        else:                                                      # :88-102
            lflag = (rX, rY, rZ)
        rflag = (rX, rY, rZ)
        ug = self.use_glassman

        r = [None] * 3
        for d in range(3):
            if not rflag[d]:
                continue
            a = _fft(self.rhs[d][0], 1, False, ug)                 # x-pass :117-158
            b = np.zeros((My, Mx), dtype=np.complex128)            # y zero-padding :226-243
            b[:ny] = a
            r[d] = _fft(b, 0, False, ug)                           # y fft :261
        if self.reg > 0:                                           # :273-287
            r = [None if v is None else v * zeta for v in r]
        kx = ikx[None, :]
        ky = iky[:, None]
        lk = [None] * 3
        if self.lhs_grad:                                          # :292-300
            lk[0] = kx * G * r[0]
            lk[1] = ky * G * r[0]
        elif self.lhs_div:                                         # :301-305
            lk[0] = G * (kx * r[0] + ky * r[1])
        elif self.lhs_curl:                                        # :306-310
            lk[0] = ky * G * r[2]
            lk[1] = -kx * G * r[2]
        else:                                                      # :311-326
            lk = [G * r[d] if lflag[d] else None for d in range(3)]

        def inverse(f):
            f = _fft(f, 0, True, ug)[:ny] / float(My)              # :331-367
            return (_fft(f, 1, True, ug) / float(Mx))[None]        # x-inverse :381-471

        for d in range(3):
            if rflag[d]:
                self.rhs[d] = inverse(r[d])
            if lflag[d]:
                self.lhs[d] = inverse(lk[d])

    # fortran/source/poisson_solver/poisson_solver_smooth3d.f:8-401: every
    # held rhs component is replaced by Re(IFFT(exp(-h^2 |k|^2 / 2) FFT(rhs)))
    # on the x-, y- and z-extended grid (y zero-filled by the map, z padded
    # in the z pass), truncated back to the physical cells
    def smooth3d(self, h):
        g = self.greens
        nx, ny, nz = self.ncell
        Mx, My = self.Mx, self.My
        ug = self.use_glassman
        k2 = (g.ikx.imag[None, None, :] ** 2 + g.iky.imag[None, :, None] ** 2
              + g.ikz.imag[:, None, None] ** 2)
        filt = np.exp(-0.5 * h * h * k2)                           # :187, :236-240
        for d in range(3):
            if self.rhs[d] is None:
                continue
            a = _fft(self.rhs[d], 2, False, ug)                    # x :77-121
            b = np.zeros((nz, My, Mx), dtype=np.complex128)        # map :123, y fft :124-177
            b[:, :ny, :] = a
            b = _fft(b, 1, False, ug)
            nfft = 2 * nz if self.bc[2] == 0 else nz               # z pad + fft :179-224
            c = np.zeros((nfft, My, Mx), dtype=np.complex128)
            c[:nz] = b
            c = _fft(c, 0, False, ug) * filt
            c = _fft(c, 0, True, ug)[:nz] / float(nfft)            # :256-270
            c = _fft(c, 1, True, ug)[:, :ny, :] / float(My)        # :272-332
            self.rhs[d] = _fft(c, 2, True, ug) / float(Mx)         # :333-386

    # settings.cpp:9-30
    def set_return_grad(self, t): self.lhs_grad = bool(t)
    def set_return_div(self, t): self.lhs_div = bool(t)
    def set_return_curl(self, t): self.lhs_curl = bool(t)
    def set_reprojection(self, t): self.rhs_reproject = bool(t)

    def _shape(self):
        nx, ny, nz = self.ncell
        return (nz, ny, nx)

    # push.cpp:9-397: real host arrays -> complex x-pencils, x zero-padded
    def push(self, bx, by=None, bz=None):
        nx, ny, nz = self.ncell
        for d, b in enumerate((bx, by, bz)):
            if b is None:
                continue
            a = np.zeros((nz, ny, self.Mx), dtype=np.complex128)
            a[:, :, :nx] = np.asarray(b, dtype=np.float64).reshape(nz, ny, nx)
            self.rhs[d] = a

    # solve3d.cpp:8-635
    def solve3d(self):
        g = self.greens
        nx, ny, nz = self.ncell
        Mx, My = self.Mx, self.My
        rX, rY, rZ = (r is not None for r in self.rhs)
        lX = lY = lZ = False
        if self.lhs_grad:                                          # :47-62
            if not rX:
                raise ValueError("grad: rhs fields have not been pushed")
            if not (rY and rZ):
                raise ValueError("grad: the reference reads rhsY and rhsZ too")
            lX = lY = lZ = True
        elif self.lhs_div:                                         # :63-76
            if not (rX and rY and rZ):
                raise ValueError("div: rhs fields have not been pushed")
            lX = True
        elif self.lhs_curl:                                        # :77-92
            if not (rX and rY and rZ):
                raise ValueError("curl: rhs fields have not been pushed")
            lX = lY = lZ = True
        else:                                                      # :93-107
            lX, lY, lZ = rX, rY, rZ
        rflag = (rX, rY, rZ)
        lflag = (lX, lY, lZ)
        ug = self.use_glassman

        r = [None] * 3
        for d in range(3):
            if not rflag[d]:
                continue
            a = _fft(self.rhs[d], 2, False, ug)                    # x-pass :124-173
            b = np.zeros((nz, My, Mx), dtype=np.complex128)        # map + y zero-fill
            b[:, :ny, :] = a
            b = _fft(b, 1, False, ug)                              # y-pass :189-237
            nfft = 2 * nz if self.bc[2] == 0 else nz               # :243-250
            c = np.zeros((nfft, My, Mx), dtype=np.complex128)
            c[:nz] = b
            r[d] = _fft(c, 0, False, ug)                           # z fft :317

        if self.reg > 0:                                           # :329-343
            for d in range(3):
                if rflag[d]:
                    r[d] = r[d] * g.zeta
        G = g.g3d
        ikx = g.ikx[None, None, :]
        iky = g.iky[None, :, None]
        ikz = g.ikz[:, None, None]
        lk = [None] * 3
        if self.lhs_grad:                                          # :348-353
            lk = [ikx * G * r[0], iky * G * r[1], ikz * G * r[2]]
        elif self.lhs_div:                                         # :354-359
            lk[0] = G * (ikx * r[0] + iky * r[1] + ikz * r[2])
        elif self.lhs_curl:                                        # :360-368
            lk = [G * (iky * r[2] - ikz * r[1]),
                  G * (ikz * r[0] - ikx * r[2]),
                  G * (ikx * r[1] - iky * r[0])]
        elif getattr(self, "lhs_lap", False):
            # not a reference option: the Laplacian of A = G * r_X, i.e. div of
            # the (component-wise) grad of one scalar, (ik . ik) G r
            lk[0] = (ikx * ikx + iky * iky + ikz * ikz) * G * r[0]
            lflag = (True, False, False)
        else:                                                      # :369-383
            lk = [G * r[d] if lflag[d] else None for d in range(3)]
        if self.rhs_reproject and rX and rY and rZ:                # :388-396
            div_psi = G * (ikx * r[0] + iky * r[1] + ikz * r[2])
            r = [r[0] + ikx * div_psi, r[1] + iky * div_psi, r[2] + ikz * div_psi]

        def inverse(f):
            nfft = f.shape[0]
            f = _fft(f, 0, True, ug)[:nz] / float(nfft)            # :403-439
            f = _fft(f, 1, True, ug) / float(My)                   # y-inverse :447-530
            f = f[:, :ny, :]                                       # map ypen2xpen
            return _fft(f, 2, True, ug) / float(Mx)                # x-inverse :541-629

        for d in range(3):
            if rflag[d]:
                self.rhs[d] = inverse(r[d])
            if lflag[d]:
                self.lhs[d] = inverse(lk[d])

    # pull.cpp:9-469: real part of x-pencils, x < nx; selected arrays zeroed first
    def pull(self, ax=None, ay=None, az=None, bx=None, by=None, bz=None):
        nx, ny, nz = self.ncell
        out = {}
        for name, user, src in (("ax", ax, self.lhs[0]), ("ay", ay, self.lhs[1]),
                                ("az", az, self.lhs[2]), ("bx", bx, self.rhs[0]),
                                ("by", by, self.rhs[1]), ("bz", bz, self.rhs[2])):
            if user is None or src is None:
                continue
            v = user.reshape(nz, ny, nx)
            v[...] = 0.0
            v[...] = src[:, :, :nx].real
            out[name] = v
        return out


def solve(ncell, bc, dx, reg_order, rhs, op="plain", reproject=False, use_glassman=False):
    """One-shot helper: setup, push ``rhs`` (list of up to 3 arrays or None),
    solve with ``op`` in {plain, grad, div, curl} (or "lap": the Laplacian of
    the potential of rhs X, an extension with no reference option), return
    (lhs list, rhs list)."""
    s = OracleSolver(use_glassman)
    s.setup3d(ncell, bc, dx, reg_order)
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    s.lhs_lap = (op == "lap")
    s.set_reprojection(reproject)
    rhs = list(rhs) + [None] * (3 - len(rhs))
    s.push(*rhs)
    s.solve3d()
    nx, ny, nz = s.ncell
    lhs = [None if s.lhs[d] is None else s.lhs[d][:, :, :nx].real.copy() for d in range(3)]
    rb = [None if s.rhs[d] is None else s.rhs[d][:, :, :nx].real.copy() for d in range(3)]
    return lhs, rb


def solve2d(ncell, bc, dx, reg_order, rhs, op="plain", use_glassman=False):
    """One-shot 2D helper: setup2d, push ``rhs`` (list of up to 3 arrays of
    shape (ny, nx) or None), solve2d with ``op`` in {plain, grad, div, curl},
    return (lhs list, rhs list) of (ny, nx) arrays."""
    s = OracleSolver(use_glassman)
    s.setup2d(ncell, bc, dx, reg_order)
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    rhs = list(rhs) + [None] * (3 - len(rhs))
    s.push(*rhs)
    s.solve2d()
    nx = s.ncell[0]
    lhs = [None if s.lhs[d] is None else s.lhs[d][0, :, :nx].real.copy() for d in range(3)]
    rb = [None if s.rhs[d] is None else s.rhs[d][0, :, :nx].real.copy() for d in range(3)]
    return lhs, rb


def rel_l2(a, b):
    """sqrt(sum (a-b)^2 / sum b^2) over all components (the reference tests'
    error norm, c++/test/unbounded3d/sources/main.cpp:252-314)."""
    if isinstance(a, (list, tuple)):
        num = sum(float(np.sum((np.asarray(x, np.float64) - np.asarray(y, np.float64)) ** 2))
                  for x, y in zip(a, b))
        den = sum(float(np.sum(np.asarray(y, np.float64) ** 2)) for y in b)
    else:
        num = float(np.sum((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
        den = float(np.sum(np.asarray(b, np.float64) ** 2))
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))
