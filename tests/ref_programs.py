"""Sources and analytic solutions of the reference's 3D validation programs
(c++/test/<name>/sources/main.cpp), on [-1,1]^3 at cell centres, arrays
(nz, ny, nx).  Each returns (rhs X/Y/Z, analytic lhs X/Y/Z, h)."""
import numpy as np

# name: (fields, bc, operator)
PROGRAMS = {}


def _grid(n):
    h = 2.0 / n
    x = -1.0 + (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return X, Y, Z, h


def mixed_ffp3d(n, c=10.0, r0=1.0):
    """x, y unbounded, z periodic; grad of a z-independent bump
    (mixed_ffp3d/sources/main.cpp:34-43, :152-190 source, :234-283 solution
    and error; the solution is set to 0 within 0.25 h of the axis, :248-253)."""
    X, Y, _, h = _grid(n)
    r2 = X * X + Y * Y
    inside = np.sqrt(r2) < r0
    d = np.where(inside, r0 ** 2 - r2, 1.0)
    e = np.where(inside, np.exp(-c * r0 ** 2 / d), 0.0)
    B = np.where(inside, 4.0 * c * r0 ** 2 * e
                 * (r0 ** 4 - X ** 4 - Y ** 4 - 2.0 * X ** 2 * Y ** 2 - c * X ** 2 * r0 ** 2 - c * Y ** 2 * r0 ** 2)
                 / d ** 4, 0.0)
    axis = np.sqrt(r2) < 0.25 * h
    sx = np.where(inside & ~axis, -2.0 * c * r0 ** 2 * X * e / d ** 2, 0.0)
    sy = np.where(inside & ~axis, -2.0 * c * r0 ** 2 * Y * e / d ** 2, 0.0)
    return [B, B.copy(), B.copy()], [sx, sy, np.zeros_like(B)], h


def mixed_ppf3d(n, c=10.0):
    """x, y periodic, z unbounded; curl of (psi_x, 0, 0) with psi_x a
    sin(pi x) sin(pi y) times a z bump (mixed_ppf3d/sources/main.cpp:34-43,
    :153-195 source, :238-279 solution and error)."""
    X, Y, Z, h = _grid(n)
    pi = np.pi
    inz = np.abs(Z) < 1.0
    z = np.where(inz, Z, 0.0)
    ez = np.where(inz, np.exp(c * z ** 2 / ((z - 1.0) * (z + 1.0))), 0.0)
    poly = (2.0 * pi ** 2 * z ** 8 - 8.0 * pi ** 2 * z ** 6 + 12.0 * pi ** 2 * z ** 4 - 6.0 * c * z ** 4
            - 8.0 * pi ** 2 * z ** 2 - 4.0 * c ** 2 * z ** 2 + 4.0 * c * z ** 2 + 2.0 * pi ** 2 + 2.0 * c)
    sxy = np.sin(pi * X) * np.sin(pi * Y)
    B = np.where(inz, ez * sxy * poly / ((z - 1.0) ** 4 * (z + 1.0) ** 4), 0.0)
    sy = np.where(inz, -2.0 * c * z * ez * sxy / ((z - 1.0) ** 2 * (z + 1.0) ** 2), 0.0)
    sz = np.where(inz, -pi * ez * np.sin(pi * X) * np.cos(pi * Y), 0.0)
    zero = np.zeros_like(B)
    return [B, zero, zero.copy()], [zero.copy(), sy, sz], h


def periodic3d(n):
    """Fully periodic; curl of a single trigonometric mode
    (periodic3d/sources/main.cpp:32-42, :148-163 source, :212-240 solution and error)."""
    X, Y, Z, h = _grid(n)
    t = 2.0 * np.pi
    B = [-np.sin(t * X) * np.cos(t * Y) * np.sin(t * Z),
         -np.cos(t * X) * np.sin(t * Y) * np.sin(t * Z),
         -2.0 * np.cos(t * X) * np.cos(t * Y) * np.cos(t * Z)]
    s = [np.cos(t * X) * np.sin(t * Y) * np.cos(t * Z) / t,
         -np.sin(t * X) * np.cos(t * Y) * np.cos(t * Z) / t,
         np.zeros_like(X)]
    return B, s, h


PROGRAMS["mixed_ffp3d"] = (mixed_ffp3d, (0, 0, 1), "grad")
PROGRAMS["mixed_ppf3d"] = (mixed_ppf3d, (1, 1, 0), "curl")
PROGRAMS["periodic3d"] = (periodic3d, (1, 1, 1), "curl")


def analytic_error(lhs, sol):
    """The programs' error integral: sqrt(sum (A - sol)^2 / sum sol^2) over all
    three components (the cell volume cancels)."""
    num = sum(float(np.sum((a - s) ** 2)) for a, s in zip(lhs, sol))
    den = sum(float(np.sum(s ** 2)) for s in sol)
    return float(np.sqrt(num / den))
