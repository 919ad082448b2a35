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


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself links
    (hipMalloc / hipMemcpy through ctypes), for device-pointer calls."""
    _hip = None

    @classmethod
    def hip(cls):
        import ctypes
        if cls._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        import ctypes
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes == self.nbytes
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1) == 0   # host to device

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            self.hip().hipFree(self.ptr)
            self.ptr.value = None
