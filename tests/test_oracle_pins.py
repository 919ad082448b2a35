"""Pins the oracle's 3D pipeline (oracle/oracle.py) to mathematics it does not
restate: brute-force lattice sums, closed-form solutions, library routines and
invariants.  Each check is chosen so that a plausible slip (a dropped weight,
a wrong sign or index, a transposed operand, a wrong normalisation or shift)
fails it."""
import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O


def cell_centres(n, lo, hi):
    h = (hi - lo) / n
    return lo + (np.arange(n) + 0.5) * h, h


# ---------------------------------------------------------------- UUU -------
@pytest.mark.parametrize("n", [(5, 4, 6), (6, 6, 6)])
def test_uuu_equals_brute_force_lattice_sum(n):
    """Doubling + kernel sampling + fftshift + truncation + normalisation:
    the FFT result must equal A_i = sum_j g(x_i - x_j) B_j with g the
    sampled regularised kernel (reference greens3d.cpp:193-225)."""
    rng = np.random.default_rng(11)
    dx = (0.11, 0.09, 0.13)
    nx, ny, nz = n
    B = rng.standard_normal((nz, ny, nx))
    lhs, _ = O.solve(n, (0, 0, 0), dx, 0, [B])
    h = max(dx)
    sigma = h / np.pi
    I = np.stack(np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij"), -1).reshape(-1, 3)
    off = (I[:, None, :] - I[None, :, :]) * np.array([dx[2], dx[1], dx[0]])
    r = np.sqrt((off ** 2).sum(-1))
    dv = dx[0] * dx[1] * dx[2]
    g = np.where(r > 0.25 * dx[0],
                 sp.sici(np.maximum(r, 1e-300) / sigma)[0] / (2 * np.pi ** 2 * np.maximum(r, 1e-300)),
                 1.0 / (2 * np.pi ** 2 * sigma)) * dv
    A = (g @ B.reshape(-1)).reshape(B.shape)
    # scipy's Si differs from NR cisi (eps 1e-10) by <= ~2e-11 relative
    assert O.rel_l2(lhs[0], A) < 1e-10


def test_uuu_gaussian_closed_form_reg0_and_reg2():
    """Config-1 source: A = erf(r/(sqrt2 s))/(4 pi r); with reg order 2 the
    filter is a Gaussian of width 2h, so s -> sqrt(s^2 + 4h^2)."""
    n = 32
    x, h = cell_centres(n, -1.0, 1.0)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    s = 0.1
    B = np.exp(-r ** 2 / (2 * s * s)) / ((2 * np.pi) ** 1.5 * s ** 3)
    for reg, se, tol in ((0, s, 1e-7), (2, np.sqrt(s * s + 4 * h * h), 1e-11)):
        lhs, rb = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, reg, [B])
        A = sp.erf(r / (np.sqrt(2) * se)) / (4 * np.pi * r)
        assert O.rel_l2(lhs[0], A) < tol, reg
    # the returned rhs is the filtered source: a Gaussian of width se
    Bf = np.exp(-r ** 2 / (2 * se * se)) / ((2 * np.pi) ** 1.5 * se ** 3)
    assert O.rel_l2(rb[0], Bf) < 1e-10


def test_uuu_kernel_limits():
    """Si(r/s)/(2 pi^2 r) -> 1/(4 pi r) for r >> s and -> 1/(2 pi^2 s) at r -> 0."""
    s = 0.01
    r = np.array([50.0, 500.0])
    g = O.sine_int(r / s) / (2 * np.pi ** 2 * r)
    assert np.allclose(g, 1 / (4 * np.pi * r), rtol=2e-4)
    r0 = np.array([1e-9])
    assert np.allclose(O.sine_int(r0 / s) / (2 * np.pi ** 2 * r0), 1 / (2 * np.pi ** 2 * s), rtol=1e-12)


def test_zeta_is_regularised_upper_incomplete_gamma():
    """zeta_m(k) = e^-a sum_{n<m/2} a^n/n! = Q(m/2, a), a = (2h)^2 k^2/2."""
    k = np.linspace(-40, 40, 81)
    h = 0.05
    for m, mm in ((2, 2), (3, 4), (4, 4), (6, 6), (10, 10)):
        z = O.zeta_filter(k, np.zeros(1), np.zeros(1), h, m)[0, 0, :]
        a = 0.5 * (2 * h) ** 2 * k ** 2
        assert np.allclose(z, sp.gammaincc(mm / 2, a), rtol=1e-13, atol=1e-15)
    assert O.zeta_filter(k, k, k, h, 0) is None


def test_ik_vector_nyquist_negative():
    ik = O.ik_vector(8, 0.5)
    m = np.array([0, 1, 2, 3, -4, -3, -2, -1])
    assert np.allclose(ik.imag, 2 * np.pi * m / (8 * 0.5), rtol=1e-15)
    ik = O.ik_vector(7, 1.0)
    m = np.array([0, 1, 2, 3, -3, -2, -1])
    assert np.allclose(ik.imag, 2 * np.pi * m / 7, rtol=1e-15)


# ---------------------------------------------------------------- grad/div/curl
def _gaussian_setup(n=32, s=0.12):
    x, h = cell_centres(n, -1.0, 1.0)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    B = np.exp(-r ** 2 / (2 * s * s)) / ((2 * np.pi) ** 1.5 * s ** 3)
    A = sp.erf(r / (np.sqrt(2) * s)) / (4 * np.pi * r)
    dA = (np.exp(-r ** 2 / (2 * s * s)) / ((2 * np.pi) ** 1.5 * s) / r
          - sp.erf(r / (np.sqrt(2) * s)) / (4 * np.pi * r * r))
    grad = [dA * X / r, dA * Y / r, dA * Z / r]
    return n, h, X, Y, Z, B, A, grad


def test_grad_div_curl_signs_and_components():
    n, h, X, Y, Z, B, A, grad = _gaussian_setup()
    lhs, _ = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, 0, [B, B, B], op="grad")
    assert O.rel_l2(lhs, grad) < 1e-6
    lhs, _ = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, 0, [B, 2 * B, -B], op="div")
    assert O.rel_l2(lhs[0], grad[0] + 2 * grad[1] - grad[2]) < 1e-6
    assert lhs[1] is None and lhs[2] is None
    # curl of (0, 0, A) = (dA/dy, -dA/dx, 0); curl of (A, 0, 0) = (0, dA/dz, -dA/dy)
    zero = np.zeros_like(B)
    lhs, _ = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, 0, [zero, zero, B], op="curl")
    assert O.rel_l2(lhs, [grad[1], -grad[0], zero]) < 1e-6
    lhs, _ = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, 0, [B, zero, zero], op="curl")
    assert O.rel_l2(lhs, [zero, grad[2], -grad[1]]) < 1e-6


def test_vortex_ring_curl_reference_case():
    """The reference's unbounded3d test (c++/test/unbounded3d/sources/main.cpp:34-44,
    :150-211, :252-314) at half resolution: curl of the solution for a compact
    vortex ring equals the analytic velocity to discretisation accuracy."""
    c, r0 = 10.0, 0.5
    n = (32, 64, 64)
    x, hx = cell_centres(n[0], -0.5, 0.5)
    y, hy = cell_centres(n[1], -1.0, 1.0)
    z, hz = cell_centres(n[2], -1.0, 1.0)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    rho = np.sqrt(Z * Z + Y * Y)
    phi = np.sqrt((rho - r0) ** 2 + X ** 2)
    th = np.arctan2(Z, Y)
    inside = phi < r0
    q = np.where(inside, 2 * r0 * rho - rho ** 2 - X ** 2, 1.0)
    e = np.where(inside, np.exp(-c * r0 ** 2 / q), 0.0)
    poly = (4 * c ** 2 * r0 ** 4 * X ** 2 * rho ** 2 - 16 * r0 ** 4 * rho ** 4 + 32 * r0 ** 3 * rho ** 5
            - 24 * r0 ** 2 * rho ** 6 + 8 * r0 * rho ** 7 - 4 * rho ** 6 * X ** 2 - 6 * rho ** 4 * X ** 4
            - 4 * rho ** 2 * X ** 6 - 8 * c * r0 ** 5 * rho ** 3 + 8 * c * r0 ** 4 * rho ** 4
            - 6 * c * r0 ** 3 * rho ** 5 + 4 * c ** 2 * r0 ** 6 * rho ** 2 - 8 * c ** 2 * r0 ** 5 * rho ** 3
            + 4 * c ** 2 * r0 ** 4 * rho ** 4 + 2 * c * r0 ** 2 * rho ** 6 + 32 * r0 ** 3 * rho ** 3 * X ** 2
            - 48 * r0 ** 2 * rho ** 4 * X ** 2 - 24 * r0 ** 2 * rho ** 2 * X ** 4 + 24 * r0 * rho ** 5 * X ** 2
            + 24 * r0 * rho ** 3 * X ** 4 + 8 * r0 * rho * X ** 6 + 2 * c * r0 ** 3 * rho * X ** 4
            + 2 * c * r0 ** 2 * rho ** 2 * X ** 4 - 4 * c * r0 ** 3 * rho ** 3 * X ** 2
            + 4 * c * r0 ** 2 * rho ** 4 * X ** 2 - rho ** 8 - X ** 8)
    Bmag = np.where(inside, -e * poly * q ** -4.0 / rho ** 2, 0.0)
    Bx, By, Bz = np.zeros_like(X), -np.sin(th) * Bmag, np.cos(th) * Bmag
    Amag = np.where(inside, 2 * c * r0 ** 2 * X * e * q ** -2.0, 0.0)
    solX = np.where(inside, e * (4 * r0 ** 2 * rho ** 2 - 4 * r0 * rho ** 3 - 4 * r0 * rho * X ** 2
                                 + rho ** 4 + 2 * rho ** 2 * X ** 2 + X ** 4 + 2 * c * r0 ** 3 * rho
                                 - 2 * c * r0 ** 2 * rho ** 2) * q ** -2.0 / rho, 0.0)
    sol = [solX, np.cos(th) * Amag, np.sin(th) * Amag]
    lhs, _ = O.solve(n, (0, 0, 0), (hx, hy, hz), 0, [Bx, By, Bz], op="curl")
    err = O.rel_l2(lhs, sol)
    assert err < 5e-3, err


# ---------------------------------------------------------------- PPP -------
def test_ppp_single_modes_exact():
    n = (8, 6, 10)
    L = (1.0, 0.75, 1.25)
    dx = tuple(L[d] / n[d] for d in range(3))
    x = np.arange(n[0]) * dx[0]
    y = np.arange(n[1]) * dx[1]
    z = np.arange(n[2]) * dx[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    a, b, c = 2, 1, 3
    kx, ky, kz = 2 * np.pi * a / L[0], 2 * np.pi * b / L[1], 2 * np.pi * c / L[2]
    B = np.sin(kx * X) * np.cos(ky * Y) * np.sin(kz * Z) + 0.7
    A = np.sin(kx * X) * np.cos(ky * Y) * np.sin(kz * Z) / (kx * kx + ky * ky + kz * kz)
    lhs, _ = O.solve(n, (1, 1, 1), dx, 0, [B])
    assert O.rel_l2(lhs[0], A) < 1e-13   # the mean (k = 0 mode) is removed
    # reprojection makes the returned rhs exactly divergence-free
    rng = np.random.default_rng(3)
    V = [rng.standard_normal(B.shape) for _ in range(3)]
    s = O.OracleSolver()
    s.setup3d(n, (1, 1, 1), dx, 0)
    s.set_reprojection(True)
    s.push(*V)
    s.solve3d()
    g = s.greens
    R = [np.fft.fftn(s.rhs[d][:, :, :n[0]].real) for d in range(3)]
    div = g.ikx[None, None, :] * R[0] + g.iky[None, :, None] * R[1] + g.ikz[:, None, None] * R[2]
    # Nyquist planes aside, ik.r = 0 after reprojection
    ok = np.ones(div.shape, bool)
    for ax, nn in ((2, n[0]), (1, n[1]), (0, n[2])):
        if nn % 2 == 0:
            idx = [slice(None)] * 3
            idx[ax] = nn // 2
            ok[tuple(idx)] = False
    assert np.abs(div[ok]).max() < 1e-10 * np.abs(R[0]).max()


# ---------------------------------------------------------------- PPU -------
def test_ppu_closed_forms():
    """x, y periodic on [0,1)^2, z unbounded on [-1,1].  The (kx,ky) = (0,0)
    column must give the 1D free-space solution -(1/2) int |z-z'| B dz'
    (intended 1D kernel, greens2d.cpp:319-338); a mode with |k_perp| > 0 must
    give the screened solution int exp(-k|z-z'|)/(2k) B dz'."""
    n = (16, 8, 64)
    dx = (1.0 / 16, 1.0 / 8, 2.0 / 64)
    x = np.arange(n[0]) * dx[0]
    y = np.arange(n[1]) * dx[1]
    z, _ = cell_centres(n[2], -1.0, 1.0)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    s = 0.08
    f = np.exp(-Z ** 2 / (2 * s * s)) / (np.sqrt(2 * np.pi) * s)
    lhs, _ = O.solve(n, (1, 1, 0), dx, 0, [f])
    A = -0.5 * (Z * sp.erf(Z / (np.sqrt(2) * s)) + s * np.sqrt(2 / np.pi) * np.exp(-Z ** 2 / (2 * s * s)))
    assert O.rel_l2(lhs[0], A) < 1e-6
    k = 2 * np.pi * np.hypot(1, 2)
    B = np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y) * f
    lhs, _ = O.solve(n, (1, 1, 0), dx, 0, [B])
    prof = (np.exp(k * k * s * s / 2) / (4 * k)
            * (np.exp(-k * Z) * sp.erfc((k * s * s - Z) / (np.sqrt(2) * s))
               + np.exp(k * Z) * sp.erfc((k * s * s + Z) / (np.sqrt(2) * s))))
    A = np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y) * prof
    assert O.rel_l2(lhs[0], A) < 1e-6


# ---------------------------------------------------------------- UUP -------
def test_uup_closed_form_2d_gaussian():
    """x, y unbounded on [-1,1]^2, z periodic: a z-independent 2D Gaussian gives
    A = -(1/2pi)(ln r + E1(r^2/2s^2)/2) (the kz = 0 plane uses the 2D kernel of
    greens2d.cpp:183-201); a z-mode cos(2 pi z) gives a screened 2D solution,
    checked against a Hankel-transform quadrature at a few radii."""
    n = (48, 48, 8)
    x, h = cell_centres(n[0], -1.0, 1.0)
    z = np.arange(n[2]) / n[2]
    dx = (h, h, 1.0 / n[2])
    Z, Y, X = np.meshgrid(z, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y)
    s = 0.1
    B = np.exp(-r ** 2 / (2 * s * s)) / (2 * np.pi * s * s)
    lhs, _ = O.solve(n, (0, 0, 1), dx, 0, [B])
    A = -(np.log(r) + 0.5 * sp.exp1(r ** 2 / (2 * s * s))) / (2 * np.pi)
    assert O.rel_l2(lhs[0], A) < 1e-7
    kz = 2 * np.pi
    lhs, _ = O.solve(n, (0, 0, 1), dx, 0, [B * np.cos(kz * Z)])
    from scipy.integrate import quad
    for (j, i) in ((24, 24), (24, 30), (20, 35)):
        rr = np.hypot(x[i], x[j])
        val = quad(lambda k: k * sp.j0(k * rr) * np.exp(-k * k * s * s / 2) / (k * k + kz * kz) / (2 * np.pi),
                   0, 200, limit=400)[0]
        assert abs(lhs[0][0, j, i] - val) < 1e-8 * max(1.0, abs(val)), (i, j)


def test_bad_boundary_tuple_rejected():
    with pytest.raises(ValueError):
        O.greens3d((4, 4, 4), (0, 1, 0), (0.1,) * 3, 0)


# ------------------------------------------------------------- smooth3d ----
def test_smooth3d_gaussian_and_single_mode():
    """smooth3d(h) (fortran poisson_solver_smooth3d.f:187, :236-240) multiplies
    the spectrum by exp(-h^2 |k|^2 / 2), the transform of a unit Gaussian of
    standard deviation h: an unbounded Gaussian of width s becomes the
    Gaussian of width sqrt(s^2 + h^2) with the same mass; a periodic single
    mode is scaled by exp(-h^2 k^2 / 2)."""
    n = 48
    x, dx = cell_centres(n, -1.0, 1.0)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r2 = X * X + Y * Y + Z * Z
    s, h = 0.08, 0.06
    g = lambda w: np.exp(-r2 / (2 * w * w)) / ((2 * np.pi) ** 1.5 * w ** 3)
    o = O.OracleSolver()
    o.setup3d((n, n, n), (0, 0, 0), (dx, dx, dx), 0)
    o.push(g(s))
    o.smooth3d(h)
    got = o.rhs[0][:, :, :n].real
    assert O.rel_l2(got, g(np.hypot(s, h))) < 1e-9
    # x, y periodic, z unbounded: a mode sin(2 pi x) cos(4 pi y) times a z Gaussian
    n3 = (16, 8, 64)
    d3 = (1.0 / 16, 1.0 / 8, 2.0 / 64)
    xs = np.arange(16) / 16.0
    ys = np.arange(8) / 8.0
    zs, _ = cell_centres(64, -1.0, 1.0)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    f = np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y) * np.exp(-Z ** 2 / (2 * s * s))
    o = O.OracleSolver()
    o.setup3d(n3, (1, 1, 0), d3, 0)
    o.push(f)
    o.smooth3d(h)
    kp2 = (2 * np.pi) ** 2 * 5
    sz = np.hypot(s, h)
    want = (np.exp(-0.5 * h * h * kp2) * np.sin(2 * np.pi * X) * np.cos(4 * np.pi * Y)
            * (s / sz) * np.exp(-Z ** 2 / (2 * sz * sz)))
    assert O.rel_l2(o.rhs[0][:, :, :16].real, want) < 1e-9


def test_oracle_laplacian_of_potential_closed_form():
    """The oracle's Laplacian of the potential ((ik.ik) G r, the extension
    output of the potential solve): for the config-1 Gaussian with reg 2 it is
    minus the Gaussian of width s_eff = sqrt(s^2 + (2h)^2)."""
    import math

    import bench
    n = 32
    b, h = bench.gaussian_source(n)
    lo, _ = O.solve((n,) * 3, (0, 0, 0), (h,) * 3, 2, [b], op="lap")
    se = math.sqrt(0.1 ** 2 + (2 * h) ** 2)
    x = -1.0 + (np.arange(n) + 0.5) * h
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    g = np.exp(-r2 / (2 * se * se)) / ((2 * np.pi) ** 1.5 * se ** 3)
    assert O.rel_l2(lo[0], -g) <= 1e-10
    # a plausible slip (the Hermitian sign of ik, i.e. +|k|^2) fails it
    assert O.rel_l2(-lo[0], -g) > 1.0
