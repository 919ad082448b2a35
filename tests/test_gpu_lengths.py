"""Specialised kernels at lengths with a factor 3 or 5 (96, 192, 320, 384,
640, 768: the doubled 48, 96, 160, 192, 320, 384 cells) in each pass position, and the
pencil-per-wavefront z kernel at 64 and 128 with z blocks of 4 and 8
planes, against the oracle (FP64 rel-L2 <= 1e-12, FP32 <= 2e-6)."""
import pytest

import poissonsolver_amd as P
from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, random_rhs

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}

CASES = [
    # (ncell, bc): Mx, My, Mz
    ((48, 96, 160), (0, 0, 0)),     # 96, 192, 320
    ((160, 48, 96), (0, 0, 0)),     # 320, 96, 192
    ((192, 24, 48), (0, 0, 0)),     # 384, 48, 96
    ((96, 96, 192), (1, 1, 0)),     # 96, 96, 384
    ((96, 192, 12), (0, 0, 1)),     # 192, 384, 12
    ((16, 16, 32), (0, 0, 0)),      # z 64: wavefront kernel, 4-plane z blocks
    ((16, 24, 64), (0, 0, 0)),      # z 128: wavefront kernel, 8-plane z blocks
    ((320, 6, 8), (0, 0, 0)),       # 640, 12, 16
    ((6, 384, 8), (0, 0, 0)),       # 12, 768, 16
    ((8, 6, 320), (0, 0, 0)),       # 16, 12, 640
    ((6, 8, 384), (0, 0, 0)),       # 12, 16, 768
    ((640, 8, 6), (1, 1, 0)),       # 640 periodic, 8, 12
    ((8, 6, 768), (1, 1, 1)),       # 8, 6, 768 periodic
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("op", ["plain", "grad", "curl"])
def test_lengths_parity(case, prec, op):
    n, bc = case
    dx = (1.0 / n[0], 1.1 / n[1], 0.9 / n[2])
    announce("lengths", n, bc, prec, op)
    b = random_rhs(n, 21)
    if op == "plain":
        b = [b[0], None, None]
    lo, _ = O.solve(n, bc, dx, 6, b, op=op)
    lg, _ = gpu_solve(n, bc, dx, 6, b, op=op, prec=prec)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= TOL[prec]


def test_specialised_kernels_selected():
    """The new lengths run the specialised kernels (not the runtime-plan ones),
    and 64 / 128 z pencils the pencil-per-wavefront kernel."""
    for n, want in (((48, 96, 160), "specialised"), ((16, 16, 32), "pencil-per-wavefront"),
                    ((16, 24, 64), "pencil-per-wavefront"), ((8, 8, 192), "specialised"),
                    ((8, 8, 320), "specialised"), ((8, 8, 384), "specialised")):
        s = P.poisson_solver(P.PS_FP32)
        s.setup3d(n, (0, 0, 0), (0.1, 0.1, 0.1), 2)
        assert s.zpass_variant() == want, (n, s.zpass_variant())
        s.close()


LONG_Y3D = [
    ((6, 512, 8), (0, 0, 0)),       # y transform 1024
    ((8, 1024, 6), (0, 0, 0)),      # y transform 2048
    ((8, 1024, 6), (1, 1, 0)),      # y periodic 1024
]


@pytest.mark.parametrize("case", LONG_Y3D, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("op", ["plain", "grad", "curl"])
@pytest.mark.parametrize("twreg_y", [-1, 0])
def test_long_y_single_precision(case, prec, op, twreg_y):
    """Single-precision y transforms of 1024 and 2048: stage twiddles in
    registers (the planner's choice there, -1) and from the table (0)."""
    n, bc = case
    dx = (0.1, 1.0 / n[1], 0.12)
    announce("long y", n, bc, prec, op, "twreg_y", twreg_y)
    b = random_rhs(n, 61)
    if op == "plain":
        b = [b[0], None, None]
    s = P.poisson_solver(P.PS_FP32 if prec == "fp32" else P.PS_MIXED)
    s.set_tuning(twreg_y=twreg_y)
    s.setup3d(n, bc, dx, 2)
    assert s.last_status == 0
    lg, _ = gpu_solve(n, bc, dx, 2, b, op=op, prec=prec, solver=s)
    s.close()
    lo, _ = O.solve(n, bc, dx, 2, b, op=op)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= TOL[prec]


@pytest.mark.parametrize("op", ["plain", "grad", "curl"])
@pytest.mark.parametrize("zwave,want", [(1, "specialised"), (2, "pencil-per-wavefront")])
def test_fp64_z_1024_fused_and_wavefront(op, zwave, want):
    """FP64 z transforms of 1024: the planner's fused specialised kernel
    (zwave 1) and the pencil-per-wavefront kernel (zwave 2)."""
    n, bc, dx = (6, 8, 512), (0, 0, 0), (0.1, 0.12, 1 / 256)
    announce("fp64 z 1024", op, zwave)
    b = random_rhs(n, 67)
    if op == "plain":
        b = [b[0], None, None]
    s = P.poisson_solver(P.PS_FP64)
    s.set_tuning(zwave=zwave)
    s.setup3d(n, bc, dx, 2)
    assert s.zpass_variant() == want
    lg, _ = gpu_solve(n, bc, dx, 2, b, op=op, prec="fp64", solver=s)
    s.close()
    lo, _ = O.solve(n, bc, dx, 2, b, op=op)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= 1e-12
