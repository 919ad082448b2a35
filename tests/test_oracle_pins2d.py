"""Pins the oracle's 2D solve (oracle/oracle.py: greens2d, OracleSolver.setup2d
and solve2d) to mathematics it does not restate: the analytic fields of the
reference's own 2D test programs (c++/test/unbounded2d, periodic2d, mixed2d),
closed forms, a direct lattice sum and a library routine.  At the reference
tests' sizes and with reg_order 0 these sources are resolved to rounding, so
a slip in the kernel, its shift, the doubling, a sign, an axis or a
normalisation fails by many orders of magnitude."""
import numpy as np
import pytest
import scipy.special as sp
from scipy.integrate import quad

from oracle import oracle as O

PI = np.pi


def cell_centres(n, lo, hi):
    h = (hi - lo) / n
    return lo + (np.arange(n) + 0.5) * h, h


# unbounded2d/sources/main.cpp:33-35, :152-170, :230-249 (c = 10, r0 = 1)
def bump(X, Y, c=10.0, r0=1.0):
    r2 = X * X + Y * Y
    inside = r2 < r0 * r0
    d = np.where(inside, r0 * r0 - r2, 1.0)
    e = np.where(inside, np.exp(-c * r0 * r0 / d), 0.0)
    B = np.where(inside, 4 * c * r0 ** 2 * e * (r0 ** 4 - X ** 4 - Y ** 4 - 2 * X ** 2 * Y ** 2
                                               - c * X ** 2 * r0 ** 2 - c * Y ** 2 * r0 ** 2) * d ** -4, 0.0)
    Ax = np.where(inside, -2 * c * r0 ** 2 * X * e * d ** -2, 0.0)
    Ay = np.where(inside, -2 * c * r0 ** 2 * Y * e * d ** -2, 0.0)
    return e, B, Ax, Ay


# mixed2d/sources/main.cpp:159-177, :239-247 (x periodic, y unbounded, c = 10)
def mixed(X, Y, c=10.0):
    ins = np.abs(Y) < 1
    d = np.where(ins, (Y - 1) * (Y + 1), 1.0)
    e = np.where(ins, np.exp(c * Y ** 2 / d), 0.0)
    poly = (PI ** 2 * Y ** 8 - 4 * PI ** 2 * Y ** 6 + 6 * PI ** 2 * Y ** 4 - 6 * c * Y ** 4 - 4 * PI ** 2 * Y ** 2
            - 4 * c ** 2 * Y ** 2 + 4 * c * Y ** 2 + PI ** 2 + 2 * c)
    B = np.where(ins, e * np.sin(PI * X) * poly / ((Y - 1) ** 4 * (Y + 1) ** 4), 0.0)
    Ax = np.where(ins, PI * e * np.cos(PI * X), 0.0)
    Ay = np.where(ins, -2 * c * Y * e * np.sin(PI * X) / ((Y - 1) ** 2 * (Y + 1) ** 2), 0.0)
    return e * np.sin(PI * X), B, Ax, Ay


# ----------------------------------------------------------------- UU -------
def test_uu_reference_case_grad_plain_curl_div():
    """unbounded2d at its own size (128^2 on [-1,1]^2, grad): the bump source is
    band-limited to rounding at this resolution, so grad, plain, curl (of the
    scalar rZ) and div (of (B, B)) meet their analytic fields to 1e-13."""
    n = 128
    x, h = cell_centres(n, -1.0, 1.0)
    X, Y = np.meshgrid(x, x)
    A, B, Ax, Ay = bump(X, Y)
    lhs, _ = O.solve2d((n, n), (0, 0), (h, h), 0, [B], "grad")
    assert lhs[2] is None
    assert O.rel_l2([lhs[0], lhs[1]], [Ax, Ay]) < 1e-13
    lhs, _ = O.solve2d((n, n), (0, 0), (h, h), 0, [B])
    assert O.rel_l2(lhs[0], A) < 1e-13
    lhs, _ = O.solve2d((n, n), (0, 0), (h, h), 0, [None, None, B], "curl")
    assert O.rel_l2([lhs[0], lhs[1]], [Ay, -Ax]) < 1e-13
    lhs, _ = O.solve2d((n, n), (0, 0), (h, h), 0, [B, B], "div")
    assert lhs[1] is None and lhs[2] is None
    assert O.rel_l2(lhs[0], Ax + Ay) < 1e-13


def test_uu_regularised_convergence_order():
    """reg_order m: the error of the regularised solve falls like h^m (the
    kernel's design order); m = 2 at least 3.5x per halving, m = 6 at least 20x."""
    errs = {2: [], 6: []}
    for n in (64, 128):
        x, h = cell_centres(n, -1.0, 1.0)
        X, Y = np.meshgrid(x, x)
        _, B, Ax, Ay = bump(X, Y)
        for m in errs:
            lhs, _ = O.solve2d((n, n), (0, 0), (h, h), m, [B], "grad")
            errs[m].append(O.rel_l2([lhs[0], lhs[1]], [Ax, Ay]))
    assert errs[2][0] / errs[2][1] > 3.5
    assert errs[6][0] / errs[6][1] > 20


def test_uu_gaussian_closed_form():
    """A 2D Gaussian of unit mass gives A = -(1/2pi)(ln r + E1(r^2/2s^2)/2): this
    fixes the additive constant ln(2 sigma) - gamma of greens2d.cpp:183-201."""
    n = 96
    x, h = cell_centres(n, -1.0, 1.0)
    X, Y = np.meshgrid(x, x)
    r = np.hypot(X, Y)
    s = 0.1
    B = np.exp(-r ** 2 / (2 * s * s)) / (2 * PI * s * s)
    lhs, _ = O.solve2d((n, n), (0, 0), (h, h), 0, [B])
    A = -(np.log(r) + 0.5 * sp.exp1(r ** 2 / (2 * s * s))) / (2 * PI)
    assert O.rel_l2(lhs[0], A) < 1e-12


def test_uu_equals_brute_force_lattice_sum():
    """Doubling + sampling + fft-shift + truncation + 1/(Mx My): the solve must
    equal A_i = sum_j g(x_i - x_j) B_j with g = -(1/2pi)(Ji0(r/s) + ln 2s - gamma)
    dx dy, Ji0 by quadrature, on an anisotropic tiny grid."""
    nx, ny = 5, 4
    dx = (0.11, 0.07)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((ny, nx))
    lhs, _ = O.solve2d((nx, ny), (0, 0), dx, 0, [B])
    s = max(dx) / PI
    c2 = np.log(2 * s) - np.euler_gamma
    ji0 = {}

    def g(r):
        if r not in ji0:
            rho = r / s
            ji0[r] = quad(lambda t: (1 - sp.j0(t)) / t if t > 0 else 0.0, 0, rho, limit=200)[0] if rho > 0 else 0.0
        return -(ji0[r] + c2) * dx[0] * dx[1] / (2 * PI)

    A = np.zeros_like(B)
    for j in range(ny):
        for i in range(nx):
            A[j, i] = sum(g(round(float(np.hypot((i - p) * dx[0], (j - q) * dx[1])), 15)) * B[q, p]
                          for q in range(ny) for p in range(nx))
    assert O.rel_l2(lhs[0], A) < 1e-10


# ----------------------------------------------------------------- PP -------
def test_pp_reference_case_and_filtered_rhs():
    """periodic2d (128^2 on [0,1)^2): B = 8 pi^2 sin sin gives A = sin sin; with
    reg_order m the pulled rhs is zeta(k) B for the single mode, zeta = Q(m/2, a)
    (regularised upper incomplete gamma), a = (2h)^2 |k|^2 / 2."""
    n = 128
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    B = 8 * PI ** 2 * np.sin(2 * PI * X) * np.sin(2 * PI * Y)
    lhs, rb = O.solve2d((n, n), (1, 1), (1.0 / n, 1.0 / n), 0, [B])
    assert O.rel_l2(lhs[0], np.sin(2 * PI * X) * np.sin(2 * PI * Y)) < 1e-13
    assert O.rel_l2(rb[0], B) < 1e-13
    for m in (2, 4, 6):
        lhs, rb = O.solve2d((n, n), (1, 1), (1.0 / n, 1.0 / n), m, [B])
        a = 0.5 * (2.0 / n) ** 2 * 2 * (2 * PI) ** 2
        z = sp.gammaincc(m // 2, a)
        assert O.rel_l2(rb[0], z * B) < 1e-13
        assert O.rel_l2(lhs[0], z * np.sin(2 * PI * X) * np.sin(2 * PI * Y)) < 1e-13


def test_pp_zero_mode_removed():
    n = 16
    B = np.ones((n, n))
    lhs, _ = O.solve2d((n, n), (1, 1), (0.1, 0.1), 0, [B])
    assert np.max(np.abs(lhs[0])) < 1e-15


# ----------------------------------------------------------- PU and UP -------
def test_pu_reference_case():
    """mixed2d at its own size (128 x 256, x periodic on [-1,1], y unbounded on
    [-2,2], grad) and the plain solve of the same source."""
    nx, ny = 128, 256
    x, hx = cell_centres(nx, -1.0, 1.0)
    y, hy = cell_centres(ny, -2.0, 2.0)
    X, Y = np.meshgrid(x, y)
    A, B, Ax, Ay = mixed(X, Y)
    lhs, _ = O.solve2d((nx, ny), (1, 0), (hx, hy), 0, [B], "grad")
    assert O.rel_l2([lhs[0], lhs[1]], [Ax, Ay]) < 1e-13
    lhs, _ = O.solve2d((nx, ny), (1, 0), (hx, hy), 0, [B])
    assert O.rel_l2(lhs[0], A) < 1e-13


def test_pu_kx0_column_1d_closed_form():
    """An x-independent source uses only the kx = 0 column (the 1D Si kernel):
    a unit Gaussian in y gives -(1/2) int |y-y'| B dy' (erf closed form)."""
    nx, ny = 8, 128
    dx = (1.0 / nx, 4.0 / ny)
    y, _ = cell_centres(ny, -2.0, 2.0)
    s = 0.1
    B = np.tile((np.exp(-y ** 2 / (2 * s * s)) / (np.sqrt(2 * PI) * s))[:, None], (1, nx))
    lhs, _ = O.solve2d((nx, ny), (1, 0), dx, 0, [B])
    A = -0.5 * (y * sp.erf(y / (np.sqrt(2) * s)) + s * np.sqrt(2 / PI) * np.exp(-y ** 2 / (2 * s * s)))
    assert O.rel_l2(lhs[0], np.tile(A[:, None], (1, nx))) < 1e-12


def test_up_transposed_reference_case():
    """UP (x unbounded, y periodic): the mixed2d case with the axes swapped.
    The reference's UP branch stores every 1D-kernel sample at a stale index
    (greens2d.cpp:400) and cannot pass this; the oracle restates the intended
    index."""
    nx, ny = 256, 128
    x, hx = cell_centres(nx, -2.0, 2.0)
    y, hy = cell_centres(ny, -1.0, 1.0)
    X, Y = np.meshgrid(x, y)
    A, B, Ay, Ax = mixed(Y, X)
    lhs, _ = O.solve2d((nx, ny), (0, 1), (hx, hy), 0, [B], "grad")
    assert O.rel_l2([lhs[0], lhs[1]], [Ax, Ay]) < 1e-13
    lhs, _ = O.solve2d((nx, ny), (0, 1), (hx, hy), 0, [B])
    assert O.rel_l2(lhs[0], A) < 1e-13


# ------------------------------------------------------------ semantics -----
def test_zeta2d_is_regularised_upper_incomplete_gamma():
    h = (0.05, 0.08)
    for m in (1, 2, 4, 6, 10):
        ikx, iky, zeta, _ = O.greens2d((12, 10), (0, 0), h, m)
        a = 0.5 * (2 * max(h)) ** 2 * (ikx.imag[None, :] ** 2 + iky.imag[:, None] ** 2)
        mm = m + (m % 2)
        assert np.allclose(zeta, sp.gammaincc(mm // 2, a), rtol=1e-13, atol=1e-300)


def test_solve2d_option_rules():
    """grad needs rX only and differentiates rX in both directions
    (solve2d.cpp:292-300); div needs rX and rY; curl needs rZ; plain maps every
    pushed component."""
    n = 8
    B = np.random.default_rng(1).standard_normal((n, n))
    with pytest.raises(ValueError):
        O.solve2d((n, n), (1, 1), (0.1, 0.1), 0, [None, B], "grad")
    with pytest.raises(ValueError):
        O.solve2d((n, n), (1, 1), (0.1, 0.1), 0, [B], "div")
    with pytest.raises(ValueError):
        O.solve2d((n, n), (1, 1), (0.1, 0.1), 0, [B, B], "curl")
    lhs, _ = O.solve2d((n, n), (1, 1), (0.1, 0.1), 0, [B, 2 * B, 3 * B])
    assert np.allclose(lhs[1], 2 * lhs[0]) and np.allclose(lhs[2], 3 * lhs[0])


def test_bad_boundary_pair_rejected():
    with pytest.raises(ValueError):
        O.greens2d((4, 4), (0, 2), (0.1, 0.1), 0)
