"""HIP path vs the oracle (FP64 restatement of the reference) on the same
seeded inputs.  FP64: rel-L2 <= 1e-12 (the bar for the FP64 GPU path).
FP32: rel-L2 <= 2e-6 against the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, random_rhs

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6

CASES = [
    # (ncell, bc, dx)
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125)),
    ((12, 10, 9), (0, 0, 0), (0.1, 0.12, 0.09)),
    ((16, 12, 20), (1, 1, 0), (1 / 16, 1 / 12, 0.1)),
    ((15, 9, 8), (1, 1, 0), (1 / 15, 1 / 9, 0.125)),
    ((16, 20, 12), (0, 0, 1), (0.1, 0.1, 1 / 12)),
    ((14, 10, 7), (0, 0, 1), (0.11, 0.09, 1 / 7)),
    ((16, 12, 10), (1, 1, 1), (1 / 16, 1 / 12, 0.1)),
    ((9, 7, 5), (1, 1, 1), (1 / 9, 1 / 7, 0.2)),
]


def _compare(n, bc, dx, reg, op, reproject, prec, tol, seed):
    announce(n, bc, reg, op, reproject, prec)
    b = random_rhs(n, seed)
    if op == "plain":
        b = [b[0], None, b[2]]
    lo, ro = O.solve(n, bc, dx, reg, b, op=op, reproject=reproject)
    lg, rg = gpu_solve(n, bc, dx, reg, b, op=op, reproject=reproject, prec=prec, want_rhs=True)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    err = O.rel_l2(got, ref)
    assert err <= tol, ("lhs", err)
    for d in range(3):
        if lo[d] is None:
            assert np.isnan(lg[d]).all()   # not written
    gr = [rg[d] for d in range(3) if ro[d] is not None]
    rr = [ro[d] for d in range(3) if ro[d] is not None]
    err_r = O.rel_l2(gr, rr)
    assert err_r <= tol, ("rhs", err_r)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("op", ["plain", "grad", "div", "curl"])
def test_fp64_parity_reg0(case, op):
    n, bc, dx = case
    _compare(n, bc, dx, 0, op, False, "fp64", FP64_TOL, 7)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("reg,reproject", [(2, False), (5, True), (6, True), (10, False)])
def test_fp64_parity_regularised(case, reg, reproject):
    n, bc, dx = case
    _compare(n, bc, dx, reg, "curl", reproject, "fp64", FP64_TOL, 11)


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("op", ["plain", "curl"])
def test_fp32_parity(case, op):
    n, bc, dx = case
    _compare(n, bc, dx, 2, op, False, "fp32", FP32_TOL, 13)


def test_fp64_parity_larger_uuu():
    n = (64, 48, 40)
    dx = (2 / 64, 1.5 / 48, 1.25 / 40)
    _compare(n, (0, 0, 0), dx, 2, "plain", False, "fp64", FP64_TOL, 17)


def test_device_special_functions_match_reference_golden():
    import os
    import poissonsolver_amd as P
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "special_ref.npz"))
    announce("special functions")
    si = P.eval_special(g["si_x"], "si")
    ji = P.eval_special(g["ji0_x"], "ji0")
    # same published algorithms; device libm may differ from glibc by an ulp.
    # Ji0(x) ~ x^2/8 near 0, where the series' absolute round-off (~1e-16)
    # dominates, so the bound is absolute + relative.
    assert np.all(np.abs(si - g["si_y"]) <= 1e-13 * np.abs(g["si_y"]))
    assert np.all(np.abs(ji - g["ji0_y"]) <= 1e-15 + 1e-13 * np.abs(g["ji0_y"]))


ODD_CASES = [
    # lengths served by E = 21, 35, 42, 14, 7 and the 3-component fallback (E = 30)
    ((21, 35, 42), (1, 1, 1), (1 / 21, 1 / 35, 1 / 42), "curl"),
    ((21, 15, 10), (0, 0, 0), (0.1, 0.1, 0.1), "curl"),
    ((60, 8, 8), (0, 0, 0), (0.05, 0.2, 0.2), "curl"),
    ((14, 7, 35), (1, 1, 0), (1 / 14, 1 / 7, 0.05), "grad"),
]


@pytest.mark.parametrize("case", ODD_CASES, ids=lambda c: "%s-%s-%s" % (c[1], "x".join(map(str, c[0])), c[3]))
def test_fp64_parity_odd_radices(case):
    n, bc, dx, op = case
    _compare(n, bc, dx, 4, op, True, "fp64", FP64_TOL, 19)


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("mixed", 2e-6)])
@pytest.mark.parametrize("op", ["plain", "grad"])
@pytest.mark.parametrize("form", [(1, "two-residue"), (2, "pencil-per-wavefront")], ids=lambda f: f[1])
def test_long_z_pencil_per_wavefront(prec, tol, op, form):
    # nz = 1024 unbounded: z transforms of 2048, FP32 storage, in both forms:
    # two 1024-point residues per pencil (zwave 1, the default) or one
    # 2048-point pencil per wavefront with 32 elements per thread (zwave 2)
    import poissonsolver_amd as P
    zwave, name = form
    n, bc, dx = (8, 6, 1024), (0, 0, 0), (0.1, 0.12, 1 / 512)
    announce("long z", n, prec, op, name)
    b = random_rhs(n, 47)
    rhs = [b[0], b[1], b[2]] if op == "grad" else [b[0]]
    s = P.poisson_solver(P.PS_FP32 if prec == "fp32" else P.PS_MIXED)
    s.set_tuning(zwave=zwave, mfma=0)     # the FP32 kernels (the mixed form's own is the matrix-core one)
    s.setup3d(n, bc, dx, 2)
    assert s.zpass_variant() == name
    lg, _ = gpu_solve(n, bc, dx, 2, rhs, op=op, prec=prec, solver=s)
    s.close()
    lo, _ = O.solve(n, bc, dx, 2, rhs, op=op)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 2e-6)])
@pytest.mark.parametrize("op", ["plain", "grad", "div", "curl"])
def test_long_z_all_operators(prec, tol, op):
    # z transforms of 2048 (nz = 1024 unbounded) with every operator: the
    # three-component z pass (div, curl) at this length, and FP64
    n, bc, dx = (6, 8, 1024), (0, 0, 0), (0.11, 0.1, 1 / 512)
    announce("long z all operators", n, prec, op)
    b = random_rhs(n, 53)
    rhs = [b[0]] if op == "plain" else [b[0], b[1], b[2]]
    lg, _ = gpu_solve(n, bc, dx, 4, rhs, op=op, prec=prec)
    lo, _ = O.solve(n, bc, dx, 4, rhs, op=op)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= tol
