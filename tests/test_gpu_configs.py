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


def _solve_prec(prec, ncell, bc, dx, reg, rhs, op, nout, sub=None):
    """One GPU solve; returns the nout output components (sub-sampled)."""
    import poissonsolver_amd as P
    from tests.gpu_util import PREC
    s = P.poisson_solver(PREC[prec])
    s.setup3d(ncell, bc, dx, reg)
    assert s.last_status == 0
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    s.push(*rhs)
    s.solve3d()
    assert s.last_status == 0
    out = [np.empty_like(rhs[0]) for _ in range(nout)]
    s.pull(*out)
    assert s.last_status == 0
    s.close()
    if sub is not None:
        out = [o[sub].copy() for o in out]
    return out


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("mixed", 1e-6)])
def test_config2_ppu_full_size(prec, tol):
    """Config 2 at its own shape: 256^3, x, y periodic, z unbounded (Mz = 512),
    reg 6, A and grad A, against the FP64 GPU solve."""
    n = 256
    announce("config 2 full", n, prec)
    src, dx = bench.ppu_source(n)
    nc = (n, n, n)
    for op, rhs, nout in (("plain", [src], 1), ("grad", [src, src, src], 3)):
        ref = _solve_prec("fp64", nc, (1, 1, 0), dx, 6, rhs, op, nout)
        got = _solve_prec(prec, nc, (1, 1, 0), dx, 6, rhs, op, nout)
        err = O.rel_l2(got, ref)
        assert err <= tol, (op, err)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 1e-6)])
def test_config3_vortex_ring_reduced(prec, tol):
    """Config 3 at 64^3 against the oracle: streamfunction psi (3 components,
    plain), velocity u = curl, div psi (one component, ~0)."""
    n = 64
    announce("config 3 reduced", n, prec)
    comps, h = bench.vortex_ring_vorticity(n)
    nc, bc, dx = (n, n, n), (0, 0, 0), (h, h, h)
    lo_psi, _ = O.solve(nc, bc, dx, 2, comps, op="plain")
    lo_u, _ = O.solve(nc, bc, dx, 2, comps, op="curl")
    lo_div, _ = O.solve(nc, bc, dx, 2, comps, op="div")
    psi = _solve_prec(prec, nc, bc, dx, 2, comps, "plain", 3)
    u = _solve_prec(prec, nc, bc, dx, 2, comps, "curl", 3)
    dv = _solve_prec(prec, nc, bc, dx, 2, comps, "div", 1)
    assert O.rel_l2(psi, lo_psi) <= tol
    assert O.rel_l2(u, lo_u) <= tol
    # div psi ~ 0: its error is measured against the velocity scale
    num = float(np.sum((dv[0] - lo_div[0]) ** 2))
    assert math.sqrt(num / sum(float(np.sum(x ** 2)) for x in lo_u)) <= tol
    ratio = math.sqrt(float(np.sum(lo_div[0] ** 2)) / sum(float(np.sum(x ** 2)) for x in lo_u))
    assert ratio < 1e-2, ratio


def test_config3_vortex_ring_full_size():
    """Config 3 at 512^3: the mixed form (matrix-core z pass, three-component
    curl / div kernels) against the FP64 GPU solve; div psi stays small
    against u."""
    n = 512
    announce("config 3 full", n)
    comps, h = bench.vortex_ring_vorticity(n)
    nc, bc, dx = (n, n, n), (0, 0, 0), (h, h, h)
    sub = (slice(None, None, 2),) * 3
    for op, nout in (("plain", 3), ("curl", 3)):
        ref = _solve_prec("fp64", nc, bc, dx, 2, comps, op, nout, sub)
        got = _solve_prec("mixed", nc, bc, dx, 2, comps, op, nout, sub)
        err = O.rel_l2(got, ref)
        assert err <= 1e-6, (op, err)
        if op == "curl":
            u = ref
    ref = _solve_prec("fp64", nc, bc, dx, 2, comps, "div", 1, sub)
    got = _solve_prec("mixed", nc, bc, dx, 2, comps, "div", 1, sub)
    usc = sum(float(np.sum(x ** 2)) for x in u)
    assert math.sqrt(float(np.sum((got[0] - ref[0]) ** 2)) / usc) <= 1e-6
    assert math.sqrt(float(np.sum(ref[0] ** 2)) / usc) < 1e-3


def test_config4_batch_of_64():
    """Config 4: the seed-4 batch of 64 uneven problems in one graph launch
    (FP32 and mixed forms) against FP64 single-context solves, and the
    problems up to 64^3 plus the first 96^3 one against the oracle."""
    import poissonsolver_amd as P
    probs, srcs = bench.batch_problems(64)
    announce("config 4 batch", len(probs))
    ref = [_solve_prec("fp64", *p, [s], "plain", 1)[0] for p, s in zip(probs, srcs)]
    for code in (P.PS_FP32, P.PS_MIXED):
        b = P.batch(probs, code)
        for i, s in enumerate(srcs):
            b.push(i, s)
        for _ in range(3):
            b.solve()
        for i, (p, s) in enumerate(zip(probs, srcs)):
            got = np.empty_like(s)
            b.pull(i, got)
            assert O.rel_l2(got, ref[i]) <= 1e-6, (code, i, p[0])
        b.close()
    seen96 = False
    for i, (p, s) in enumerate(zip(probs, srcs)):
        n = p[0][0]
        if n > 96 or (n == 96 and seen96):
            continue
        seen96 = seen96 or n == 96
        lo, _ = O.solve(*p, [s])
        assert O.rel_l2(ref[i], lo[0]) <= 1e-12, (i, n)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 1e-6)])
def test_config5_blobs_reduced(prec, tol):
    """Config 5 at 64^3 against the oracle: 256 Gaussian blobs, reg 10, A and
    grad A."""
    n = 64
    announce("config 5 reduced", n, prec)
    src, h = bench.large_blobs(n)
    nc, bc, dx = (n, n, n), (0, 0, 0), (h, h, h)
    for op, rhs, nout in (("plain", [src], 1), ("grad", [src, src, src], 3)):
        lo, _ = O.solve(nc, bc, dx, 10, rhs, op=op)
        got = _solve_prec(prec, nc, bc, dx, 10, rhs, op, nout)
        assert O.rel_l2(got, lo[:nout]) <= tol, op


def test_config5_blobs_full_size():
    """Config 5 at 1024^3 (Mz = 2048): A in the mixed and FP32 forms against
    the FP64 GPU solve on the 256^3 sub-sample of every fourth cell."""
    n = 1024
    announce("config 5 full", n)
    src, h = bench.large_blobs(n)
    nc, bc, dx = (n, n, n), (0, 0, 0), (h, h, h)
    sub = (slice(None, None, 4),) * 3
    ref = _solve_prec("fp64", nc, bc, dx, 10, [src], "plain", 1, sub)
    for prec in ("mixed", "fp32"):
        got = _solve_prec(prec, nc, bc, dx, 10, [src], "plain", 1, sub)
        err = O.rel_l2(got, ref)
        assert err <= 1e-6, (prec, err)
