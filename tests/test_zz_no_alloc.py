"""ps_solve3d_noalloc (a solve whose work space is a static device array, for
callers that cannot allocate) against the allocating path of the same context
(ps_push_device + ps_solve3d_async + ps_pull_device): bit-identical outputs,
on two calls in a row, the second with a smaller input; oversize grids are
refused before anything is launched."""
import ctypes

import numpy as np
import pytest

import poissonsolver_amd as P
from poissonsolver_amd import _capi
from tests.gpu_util import PREC, DeviceBuffer, announce

pytestmark = pytest.mark.gpu


def _noalloc_fn(lib):
    fn = lib.ps_solve3d_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return fn


def _case(prec, n, bc, dx, reg, op, mask, seed):
    announce("noalloc", prec, n, bc, op, mask)
    dt = np.float64 if prec == "fp64" else np.float32
    eb = np.dtype(dt).itemsize
    n3 = n[0] * n[1] * n[2]
    rng = np.random.default_rng(seed)
    src = np.zeros((3, n3), dtype=dt)
    for d in range(3):
        if (mask >> d) & 1:
            src[d] = rng.standard_normal(n3)
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, reg)
    assert s.last_status == 0
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    d_src = DeviceBuffer(3 * n3 * eb)
    d_src.upload(src)
    d_new = DeviceBuffer(3 * n3 * eb)
    d_new.upload(np.full((3, n3), np.nan, dtype=dt))
    d_old = DeviceBuffer(3 * n3 * eb)
    d_old.upload(np.full((3, n3), np.nan, dtype=dt))
    fn = _noalloc_fn(s._lib)
    assert fn(s._ctx, d_src.address, mask, d_new.address) == 0, s._lib.ps_last_error(s._ctx)
    assert s._lib.ps_synchronize(s._ctx) == 0
    new = d_new.download(dt, (3, n3))
    # the allocating path of the same context on the same data
    ptr = [d_src.address + d * n3 * eb if (mask >> d) & 1 else None for d in range(3)]
    s.push_device(*ptr)
    s.solve3d()
    out = [d_old.address + d * n3 * eb for d in range(3)]
    s.pull_device(*out)
    assert s.last_status == 0
    old = d_old.download(dt, (3, n3))
    for d in range(3):
        assert np.array_equal(np.isnan(new[d]), np.isnan(old[d])), d
        assert np.array_equal(np.nan_to_num(new[d]), np.nan_to_num(old[d])), d
    assert np.isfinite(new).any()
    s.close()
    for b in (d_src, d_new, d_old):
        b.free()


CASES = {
    # first call, then a smaller one
    "fp64": [((40, 36, 32), (0, 0, 0), (0.05, 0.05, 0.06), 2, "plain", 1),
             ((20, 18, 16), (1, 1, 0), (0.05, 0.055, 0.06), 6, "curl", 7)],
    "fp32": [((64, 48, 40), (0, 0, 0), (0.05, 0.05, 0.06), 2, "grad", 7),
             ((24, 20, 28), (0, 0, 1), (0.05, 0.05, 0.04), 0, "plain", 5)],
    "mixed": [((16, 16, 512), (0, 0, 0), (0.1, 0.1, 1 / 256), 2, "plain", 1),   # matrix-core z pass
              ((12, 10, 512), (0, 0, 0), (0.1, 0.1, 1 / 256), 0, "div", 7)],
}


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_noalloc_matches_allocating_solve(prec):
    for k, (n, bc, dx, reg, op, mask) in enumerate(CASES[prec]):
        _case(prec, n, bc, dx, reg, op, mask, 500 + k)


def test_noalloc_refuses_oversize_and_bad_arguments():
    n = (256, 256, 256)
    s = P.poisson_solver(P.PS_FP64)
    s.setup3d(n, (0, 0, 0), (1 / 128,) * 3, 0)
    assert s.last_status == 0
    fn = _noalloc_fn(s._lib)
    buf = DeviceBuffer(8 * 3 * 1024)        # never touched: the call is refused first
    assert fn(s._ctx, buf.address, 1, buf.address) == _capi.PS_EUNSUPPORTED
    assert b"static maximum" in s._lib.ps_last_error(s._ctx)
    assert fn(s._ctx, buf.address, 0, buf.address) == _capi.PS_EINVAL
    assert fn(s._ctx, None, 1, buf.address) == _capi.PS_EINVAL
    s.close()
    buf.free()
