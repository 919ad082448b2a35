"""Edge shapes against the oracle: directions of one or two cells (length-1
and length-2 transforms), strongly anisotropic grids, and a 3D UUP grid with
one z cell, which is the 2D unbounded problem (its kz = 0 plane is the 2D
kernel of greens2d.cpp)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, gpu_solve2d, random_rhs

pytestmark = pytest.mark.gpu

CASES = [
    ((8, 8, 1), (1, 1, 1)),
    ((1, 8, 8), (1, 1, 1)),
    ((8, 1, 8), (0, 0, 0)),
    ((1, 1, 1), (0, 0, 0)),
    ((2, 3, 1), (0, 0, 1)),
    ((64, 2, 2), (0, 0, 0)),
    ((2, 2, 64), (1, 1, 0)),
    ((3, 40, 5), (1, 1, 0)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % ("x".join(map(str, c[0])), "".join(map(str, c[1]))))
@pytest.mark.parametrize("op", ["plain", "grad", "curl"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_edge_shapes(case, op, prec):
    n, bc = case
    dx = (0.1, 0.13, 0.07)
    announce("edge", n, bc, op, prec)
    b = random_rhs(n, 5)
    if op == "plain":
        b = [b[0], None, None]
    lo, _ = O.solve(n, bc, dx, 2, b, op=op)
    lg, _ = gpu_solve(n, bc, dx, 2, b, op=op, prec=prec)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    tol = 1e-12 if prec == "fp64" else 2e-6
    assert O.rel_l2(got, ref) <= tol


@pytest.mark.parametrize("n", [(1, 6), (6, 1), (1, 1), (40, 2)])
@pytest.mark.parametrize("bc", [(0, 0), (1, 1), (1, 0), (0, 1)])
def test_edge_shapes_2d(n, bc):
    announce("edge2d", n, bc)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((n[1], n[0]))
    lo, _ = O.solve2d(n, bc, (0.1, 0.12), 2, [b], op="grad")
    lg, _ = gpu_solve2d(n, bc, (0.1, 0.12), 2, [b], "grad")
    assert O.rel_l2([lg[0], lg[1]], [lo[0], lo[1]]) <= 1e-12


def test_uup_one_z_cell_is_the_2d_solve():
    """3D UUP with nz = 1 and the 2D UU solve of the same plane agree (the
    kz = 0 plane of the UUP kernel is the 2D kernel; nz = 1 has no other)."""
    n2 = (24, 20)
    dx = (0.1, 0.08)
    rng = np.random.default_rng(8)
    b = rng.standard_normal((n2[1], n2[0]))
    l3, _ = gpu_solve((n2[0], n2[1], 1), (0, 0, 1), (dx[0], dx[1], 0.05), 0, [b[None]])
    l2, _ = gpu_solve2d(n2, (0, 0), dx, 0, [b])
    assert O.rel_l2(l3[0][0], l2[0]) <= 1e-12
