"""CPU tests of the C ABI: the library loads, exports every function that
include/ps_capi.h declares and the C++ class symbols of
include/poisson_solver.hpp; host-side validation and planning work without a
GPU (no compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

import poissonsolver_amd as P
from poissonsolver_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "ps_capi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ps_\w+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for each of them
    assert set(names) <= set(_capi.SIGNATURES)


def test_cpp_class_symbols_exported():
    out = subprocess.run(["nm", "-DC", _capi.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ["poisson_solver::setup3d(int*, int*, double*, int)", "poisson_solver::solve3d()",
                "poisson_solver::push(double*, double*, double*)",
                "poisson_solver::pull(double*, double*, double*, double*, double*, double*)",
                "poisson_solver::set_return_curl(bool)", "poisson_solver::set_reprojection(bool)",
                "poisson_solver::setup2d(int*, int*, double*, int)",
                "partition_setup(int, int*, int*, double*, bool, bool, bool)"]:
        assert sym in out, sym


def test_host_validation_without_device():
    lib = _capi.lib()
    ctx = ctypes.c_void_p()
    assert lib.ps_create(ctypes.byref(ctx)) == _capi.PS_OK
    n = (ctypes.c_int * 3)(8, 8, 8)
    dx = (ctypes.c_double * 3)(0.1, 0.1, 0.1)
    # boundary tuples the reference rejects (greens3d.cpp:522-541)
    for bc in [(0, 1, 0), (1, 0, 0), (0, 1, 1), (1, 0, 1), (2, 0, 0)]:
        b = (ctypes.c_int * 3)(*bc)
        assert lib.ps_setup3d(ctx, n, b, dx, 0) == _capi.PS_EBC, bc
        if bc != (2, 0, 0):
            assert "z-direction" in lib.ps_last_error(ctx).decode()
    bad_n = (ctypes.c_int * 3)(8, 0, 8)
    b = (ctypes.c_int * 3)(0, 0, 0)
    assert lib.ps_setup3d(ctx, bad_n, b, dx, 0) == _capi.PS_EINVAL
    assert lib.ps_setup3d(ctx, n, b, dx, -1) == _capi.PS_EINVAL
    # calls out of order
    assert lib.ps_solve3d(ctx) == _capi.PS_ESTATE
    assert lib.ps_push(ctx, None, None, None) == _capi.PS_ESTATE
    assert lib.ps_set_precision(ctx, 7) == _capi.PS_EINVAL
    # setup2d: argument checks run before any device call (greens2d.cpp:430-434)
    b2 = (ctypes.c_int * 2)(0, 2)
    assert lib.ps_setup2d(ctx, n, b2, dx, 0) == _capi.PS_EBC
    assert "Boundary condition configuration unknown" in lib.ps_last_error(ctx).decode()
    n2 = (ctypes.c_int * 2)(8, 0)
    assert lib.ps_setup2d(ctx, n2, b, dx, 0) == _capi.PS_EINVAL
    assert lib.ps_setup2d(ctx, n, b, dx, -2) == _capi.PS_EINVAL
    assert lib.ps_setup2d(ctx, n, b, dx, 0) == _capi.PS_EHIP       # valid, but no device here
    assert lib.ps_solve2d(ctx) == _capi.PS_ESTATE
    assert lib.ps_is_2d(ctx) == 0
    assert lib.ps_partition_count(ctx) == 0
    assert lib.ps_destroy(ctx) == _capi.PS_OK
    assert lib.ps_destroy(None) == _capi.PS_EINVAL
    assert _capi.strerror(_capi.PS_EBC) == "unsupported boundary conditions"


def test_python_mirror_reports_like_the_reference(capsys):
    s = P.poisson_solver()
    s.setup3d((8, 8, 8), (0, 1, 0), (0.1, 0.1, 0.1))
    assert s.last_status == _capi.PS_EBC
    assert "poisson_solver.setup3d" in capsys.readouterr().err
    s.solve3d()
    assert s.last_status == _capi.PS_ESTATE


@pytest.mark.parametrize("L,expect", [
    (1024, (16, 64, [16, 16, 4])), (2048, (16, 128, [16, 16, 8])), (64, (16, 4, [16, 4])),
    (96, (12, 8, [4, 4, 3, 2])), (320, (20, 16, [5, 4, 4, 4])), (7, (7, 1, [7])), (1, (1, 1, [])),
])
def test_plan_radix_splits(L, expect):
    assert P.plan_describe(L) == expect


@pytest.mark.parametrize("L", list(range(1, 400)) + [512, 640, 768, 1000, 1024, 1536, 2048, 3000, 4096])
def test_plan_covers_7_smooth_lengths(L):
    """Every length whose prime factors are <= 7 gets a plan, except those that
    need all of 3, 5 and 7 (no compiled E holds them); others are refused."""
    m = L
    primes = set()
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
            primes.add(p)
    plan = P.plan_describe(L)
    if m != 1 or {3, 5, 7} <= primes:
        assert plan is None
        return
    assert plan is not None, L
    E, tp, radices = plan
    prod = 1
    for r in radices:
        assert E % r == 0
        prod *= r
    assert prod == L and E * tp == L and tp <= 512
    E3, tp3, _ = P.plan_describe(L, ncomp=3)
    if E <= 10:
        assert E3 <= 10
