"""Seeded sweep of random configurations against the oracle: grid sizes
(including lengths with a factor 3 or 5), boundary tuples, operators,
regularisation orders and precisions drawn from a fixed generator, in 3D
and 2D.  FP64 rel-L2 <= 1e-12, FP32 <= 2e-6, mixed <= 1e-6."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, gpu_solve2d

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}
SIZES = [4, 5, 6, 8, 9, 10, 12, 15, 16, 20, 24, 25, 27, 30, 32, 36, 40, 48]
BC3 = [(0, 0, 0), (1, 1, 0), (0, 0, 1), (1, 1, 1)]
BC2 = [(0, 0), (1, 0), (0, 1), (1, 1)]


def _draw3(i):
    r = np.random.default_rng(1000 + i)
    n = tuple(int(r.choice(SIZES)) for _ in range(3))
    bc = BC3[int(r.integers(4))]
    op = ["plain", "grad", "div", "curl"][int(r.integers(4))]
    reg = int(r.choice([0, 2, 3, 4, 6, 8]))
    reproj = bool(r.integers(2)) and op != "plain"
    prec = ["fp64", "fp64", "fp32", "mixed"][int(r.integers(4))]
    dx = tuple(float(v) for v in r.uniform(0.05, 0.2, 3))
    return n, bc, op, reg, reproj, prec, dx


def _draw2(i):
    r = np.random.default_rng(2000 + i)
    n = tuple(int(r.choice(SIZES + [64, 96])) for _ in range(2))
    bc = BC2[int(r.integers(4))]
    op = ["plain", "grad", "div", "curl"][int(r.integers(4))]
    reg = int(r.choice([0, 2, 4, 6]))
    prec = ["fp64", "fp64", "fp32", "mixed"][int(r.integers(4))]
    dx = tuple(float(v) for v in r.uniform(0.05, 0.2, 2))
    return n, bc, op, reg, prec, dx


@pytest.mark.parametrize("i", range(40))
def test_sweep_3d(i):
    n, bc, op, reg, reproj, prec, dx = _draw3(i)
    announce("sweep3d", i, n, bc, op, reg, reproj, prec)
    rng = np.random.default_rng(i)
    b = [rng.standard_normal((n[2], n[1], n[0])) for _ in range(3)]
    if op == "plain":
        b = [b[0], None, b[2]]
    lo, ro = O.solve(n, bc, dx, reg, b, op=op, reproject=reproj)
    lg, rg = gpu_solve(n, bc, dx, reg, b, op=op, reproject=reproj, prec=prec, want_rhs=True)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= TOL[prec]
    gr = [rg[d] for d in range(3) if ro[d] is not None]
    rr = [ro[d] for d in range(3) if ro[d] is not None]
    assert O.rel_l2(gr, rr) <= TOL[prec]


@pytest.mark.parametrize("i", range(30))
def test_sweep_2d(i):
    n, bc, op, reg, prec, dx = _draw2(i)
    announce("sweep2d", i, n, bc, op, reg, prec)
    rng = np.random.default_rng(i)
    b = [rng.standard_normal((n[1], n[0])) for _ in range(3)]
    b = {"plain": [b[0], None, b[2]], "grad": [b[0]], "div": [b[0], b[1]], "curl": [None, None, b[2]]}[op]
    lo, ro = O.solve2d(n, bc, dx, reg, b, op=op)
    lg, rg = gpu_solve2d(n, bc, dx, reg, b, op=op, prec=prec, want_rhs=True)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= TOL[prec]
    gr = [rg[d] for d in range(3) if ro[d] is not None]
    rr = [ro[d] for d in range(3) if ro[d] is not None]
    assert O.rel_l2(gr, rr) <= TOL[prec]
