"""The benchmark configurations at sizes the oracle finishes in seconds.

Config 1 exactly: fully unbounded 64^3 cell-centred grid on [-1,1]^3, FP64,
reg order 2, B = exp(-r^2/(2 s^2)) / ((2 pi)^{3/2} s^3), s = 0.1.  With reg 2
the filter is a Gaussian of width sigma = 2h, so the solution is
erf(r / (sqrt 2 s_eff)) / (4 pi r) with s_eff^2 = s^2 + sigma^2.
Config 2 reduced: x, y periodic / z unbounded, reg 6, the sin*sin*gauss +
random-phase source, outputs A and grad A in FP64, FP32 and mixed forms."""
import math
import os
import sys

import numpy as np
import pytest
from scipy.special import erf

from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402  (source generators of the benchmark configurations)

pytestmark = pytest.mark.gpu


def test_config1_gaussian_fp64():
    n = 64
    announce("config 1", n)
    b, h = bench.gaussian_source(n)
    lg, _ = gpu_solve((n, n, n), (0, 0, 0), (h, h, h), 2, [b], op="plain", prec="fp64")
    lo, _ = O.solve((n, n, n), (0, 0, 0), (h, h, h), 2, [b], op="plain")
    assert O.rel_l2([lg[0]], [lo[0]]) <= 1e-12
    x = -1.0 + (np.arange(n) + 0.5) * h
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    se = math.sqrt(0.1 ** 2 + (2 * h) ** 2)
    A = erf(r / (math.sqrt(2) * se)) / (4 * np.pi * r)
    err = float(np.sqrt(np.sum((lg[0] - A) ** 2) / np.sum(A ** 2)))
    assert err <= 1e-10, err


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 1e-6)])
def test_config2_ppu_reduced(prec, tol):
    n = 32
    announce("config 2 reduced", n, prec)
    src, dx = bench.ppu_source(n)
    nc = (n, n, n)
    for op, rhs in (("plain", [src]), ("grad", [src, src, src])):
        lg, _ = gpu_solve(nc, (1, 1, 0), dx, 6, rhs, op=op, prec=prec)
        lo, _ = O.solve(nc, (1, 1, 0), dx, 6, rhs, op=op)
        got = [lg[d] for d in range(3) if lo[d] is not None]
        ref = [lo[d] for d in range(3) if lo[d] is not None]
        assert O.rel_l2(got, ref) <= tol, (op, O.rel_l2(got, ref))
