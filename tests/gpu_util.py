"""Helpers shared by the GPU tests: run the HIP path through the Python mirror
of the reference API and the oracle on the same inputs."""
import sys

import numpy as np

import poissonsolver_amd as P
from oracle import oracle as O

PREC = {"fp64": P.PS_FP64, "fp32": P.PS_FP32, "mixed": P.PS_MIXED}


def announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def gpu_solve(n, bc, dx, reg, rhs, op="plain", reproject=False, prec="fp64", want_rhs=False, solver=None):
    """setup3d -> push -> solve3d -> pull on the GPU (on `solver`, already set
    up, when given; it is left open)."""
    s = solver or P.poisson_solver(PREC[prec])
    if solver is None:
        s.setup3d(n, bc, dx, reg)
        assert s.last_status == 0, s.last_status
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    s.set_reprojection(reproject)
    rhs = list(rhs) + [None] * (3 - len(rhs))
    s.push(*[None if r is None else np.ascontiguousarray(r, np.float64) for r in rhs])
    assert s.last_status == 0
    s.solve3d()
    assert s.last_status == 0, s.last_status
    shape = (n[2], n[1], n[0])
    lhs = [np.full(shape, np.nan) for _ in range(3)]
    rb = [np.full(shape, np.nan) for _ in range(3)]
    s.pull(*lhs, *(rb if want_rhs else [None] * 3))
    assert s.last_status == 0
    if solver is None:
        s.close()
    return lhs, rb


def random_rhs(n, seed, ncomp=3):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal((n[2], n[1], n[0])) for _ in range(ncomp)]


def gpu_solve2d(n, bc, dx, reg, rhs, op="plain", prec="fp64", want_rhs=False, solver=None):
    """setup2d -> push -> solve2d -> pull on the GPU; arrays of shape (ny, nx)."""
    s = solver or P.poisson_solver(PREC[prec])
    if solver is None:
        s.setup2d(n, bc, dx, reg)
        assert s.last_status == 0, s.last_status
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    rhs = list(rhs) + [None] * (3 - len(rhs))
    s.push(*[None if r is None else np.ascontiguousarray(r, np.float64) for r in rhs])
    assert s.last_status == 0
    s.solve2d()
    assert s.last_status == 0, s.last_status
    shape = (n[1], n[0])
    lhs = [np.full(shape, np.nan) for _ in range(3)]
    rb = [np.full(shape, np.nan) for _ in range(3)]
    s.pull(*lhs, *(rb if want_rhs else [None] * 3))
    assert s.last_status == 0
    if solver is None:
        s.close()
    return lhs, rb
