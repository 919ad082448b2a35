"""2D solver (setup2d -> push -> solve2d -> pull) on the GPU against the oracle's
restatement of greens2d.cpp / solve2d.cpp, on the same seeded inputs, for the
four boundary pairs and the four operators.  FP64 rel-L2 <= 1e-12, FP32
<= 2e-6, mixed <= 1e-6.  The reference's own 2D test programs run at their
sizes against their analytic fields."""
import numpy as np
import pytest

import poissonsolver_amd as P
from poissonsolver_amd import _capi
from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve2d
from tests.test_oracle_pins2d import bump, cell_centres, mixed

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}

CASES = [
    # (ncell, bc, dx)
    ((16, 16), (0, 0), (0.125, 0.125)),
    ((12, 10), (0, 0), (0.1, 0.13)),
    ((16, 20), (1, 0), (1 / 16, 0.1)),
    ((15, 9), (1, 0), (1 / 15, 0.125)),
    ((20, 16), (0, 1), (0.1, 1 / 16)),
    ((9, 14), (0, 1), (0.125, 1 / 14)),
    ((16, 12), (1, 1), (1 / 16, 1 / 12)),
    ((9, 7), (1, 1), (1 / 9, 1 / 7)),
]


def _rhs(n, seed, op):
    rng = np.random.default_rng(seed)
    b = [rng.standard_normal((n[1], n[0])) for _ in range(3)]
    if op == "plain":
        return [b[0], None, b[2]]
    if op == "grad":
        return [b[0], None, None]
    if op == "div":
        return [b[0], b[1], None]
    return [None, None, b[2]]


def _compare(n, bc, dx, reg, op, prec, seed=7):
    announce("2d", n, bc, reg, op, prec)
    b = _rhs(n, seed, op)
    lo, ro = O.solve2d(n, bc, dx, reg, b, op=op)
    lg, rg = gpu_solve2d(n, bc, dx, reg, b, op=op, prec=prec, want_rhs=True)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    err = O.rel_l2(got, ref)
    assert err <= TOL[prec], ("lhs", err)
    for d in range(3):
        if lo[d] is None:
            assert np.isnan(lg[d]).all()   # not written
    gr = [rg[d] for d in range(3) if ro[d] is not None]
    rr = [ro[d] for d in range(3) if ro[d] is not None]
    err_r = O.rel_l2(gr, rr)
    assert err_r <= TOL[prec], ("rhs", err_r)
    return err


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("op", ["plain", "grad", "div", "curl"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_2d_parity_reg0(case, op, prec):
    n, bc, dx = case
    _compare(n, bc, dx, 0, op, prec)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("reg,op", [(2, "grad"), (5, "curl"), (6, "div"), (10, "plain")])
def test_2d_parity_regularised(case, reg, op):
    n, bc, dx = case
    _compare(n, bc, dx, reg, op, "fp64", seed=11)


@pytest.mark.parametrize("n,bc", [((64, 128), (0, 0)), ((96, 64), (1, 0)), ((200, 150), (0, 1)),
                                  ((250, 120), (1, 1)), ((384, 256), (0, 0))])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_2d_parity_medium(n, bc, prec):
    """Lengths with specialised and pencil-per-wavefront kernels (256, 512) and
    runtime plans with radices 3 and 5 (150, 250, 300, 600, 768)."""
    dx = (1.0 / n[0], 1.3 / n[1])
    _compare(n, bc, dx, 6, "grad", prec, seed=3)


@pytest.mark.parametrize("op", ["plain", "grad", "curl"])
@pytest.mark.parametrize("reg", [0, 6])
def test_2d_mixed_matrix_core(op, reg):
    """ny = 512 unbounded: the y transform has length 1024 and the mixed form
    runs it on the matrix cores (plain, grad, curl and the filtered rhs)."""
    n = (96, 512)
    s = P.poisson_solver(P.PS_MIXED)
    s.setup2d(n, (0, 0), (1 / 96, 1 / 512), reg)
    assert s.zpass_variant() == "matrix-core"
    s.close()
    _compare(n, (0, 0), (1 / 96, 1 / 512), reg, op, "mixed", seed=5)


def test_2d_large_fp64():
    """2048^2 unbounded (transform lengths 4096): FP64 against the oracle."""
    n = (2048, 2048)
    _compare(n, (0, 0), (1 / 2048, 1 / 2048), 6, "grad", "fp64", seed=2)


# --------------------------------------------------- the reference's 2D tests
def test_reference_unbounded2d():
    """c++/test/unbounded2d: 128^2 on [-1,1]^2, grad of the bump."""
    n = 128
    x, h = cell_centres(n, -1.0, 1.0)
    X, Y = np.meshgrid(x, x)
    A, B, Ax, Ay = bump(X, Y)
    for prec, tol in (("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 1e-6)):
        announce("unbounded2d", prec)
        lg, _ = gpu_solve2d((n, n), (0, 0), (h, h), 0, [B], "grad", prec=prec)
        assert O.rel_l2([lg[0], lg[1]], [Ax, Ay]) < tol, prec


def test_reference_periodic2d():
    """c++/test/periodic2d: 128^2 on [0,1)^2, A = sin sin, pull of the rhs too."""
    n = 128
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    B = 8 * np.pi ** 2 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    lg, rg = gpu_solve2d((n, n), (1, 1), (1 / n, 1 / n), 0, [B], want_rhs=True)
    assert O.rel_l2(lg[0], np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)) < 1e-12
    assert O.rel_l2(rg[0], B) < 1e-12


def test_reference_mixed2d_and_transposed():
    """c++/test/mixed2d (128 x 256, PU, grad) and the same case with the axes
    swapped (UP, the branch whose reference kernel has a stale index)."""
    nx, ny = 128, 256
    x, hx = cell_centres(nx, -1.0, 1.0)
    y, hy = cell_centres(ny, -2.0, 2.0)
    X, Y = np.meshgrid(x, y)
    A, B, Ax, Ay = mixed(X, Y)
    lg, _ = gpu_solve2d((nx, ny), (1, 0), (hx, hy), 0, [B], "grad")
    assert O.rel_l2([lg[0], lg[1]], [Ax, Ay]) < 1e-12
    XT, YT = np.meshgrid(y, x)
    A, B, Ay, Ax = mixed(YT, XT)
    lg, _ = gpu_solve2d((ny, nx), (0, 1), (hy, hx), 0, [B], "grad")
    assert O.rel_l2([lg[0], lg[1]], [Ax, Ay]) < 1e-12


# ------------------------------------------------------------- semantics ----
def test_2d_state_and_errors():
    """partition after setup2d is (nx, ny, 1) with dx (dx, dy, 0)
    (setup2d.cpp:31-46); solve3d on a 2D set-up and solve2d on a 3D one fail;
    lhs components of an earlier solve persist (solve2d.cpp:196-211 allocates
    only the selected ones); option precedence grad > div > curl."""
    n = (24, 20)
    dx = (0.1, 0.05)
    rng = np.random.default_rng(9)
    b = [rng.standard_normal((n[1], n[0])) for _ in range(3)]
    s = P.poisson_solver()
    s.setup2d(n, (0, 1), dx, 2)
    assert s.last_status == 0
    assert len(s.partition) == 1
    assert s.partition[0].ncell == [24, 20, 1] and s.partition[0].dx == [0.1, 0.05, 0.0]
    s.push(b[0])
    s.solve3d()
    assert s.last_status == _capi.PS_ESTATE
    # curl needs rZ
    s.set_return_curl(True)
    s.solve2d()
    assert s.last_status == _capi.PS_ESTATE
    # grad wins over curl and div
    s.set_return_div(True)
    s.set_return_grad(True)
    s.solve2d()
    assert s.last_status == 0
    g = [np.full((n[1], n[0]), np.nan) for _ in range(3)]
    s.pull(*g)
    lo, _ = O.solve2d(n, (0, 1), dx, 2, [b[0]], op="grad")
    assert O.rel_l2([g[0], g[1]], [lo[0], lo[1]]) < 1e-12
    assert np.isnan(g[2]).all()
    # then a plain solve with rZ pushed (rX still held): lhs X and Z are
    # recomputed, lhs Y keeps the grad result (reg 0: the held rX is the pushed one)
    s2 = P.poisson_solver()
    s2.setup2d(n, (0, 1), dx, 0)
    lg, _ = gpu_solve2d(n, (0, 1), dx, 0, [b[0]], "grad", solver=s2)
    lg2, _ = gpu_solve2d(n, (0, 1), dx, 0, [None, None, b[2]], "plain", solver=s2)
    lz, _ = O.solve2d(n, (0, 1), dx, 0, [b[0], None, b[2]])
    assert O.rel_l2([lg2[0], lg2[2]], [lz[0], lz[2]]) < 1e-12
    assert np.array_equal(lg2[1], lg[1])
    s2.close()
    s3 = P.poisson_solver()
    s3.setup3d((8, 8, 8), (0, 0, 0), (0.1,) * 3)
    s3.solve2d()
    assert s3.last_status == _capi.PS_ESTATE
    s3.close()
    s.close()


def test_2d_profile_and_autotune():
    """ps_profile_solve reports x, fused y (as z) and x^-1; the y entries of a
    2D solve are empty.  ps_autotune keeps the 2D set-up."""
    n = (256, 256)
    s = P.poisson_solver(P.PS_FP32)
    s.setup2d(n, (0, 0), (1 / 256, 1 / 256), 2)
    s.push(np.random.default_rng(0).standard_normal((256, 256)))
    t = s.profile_solve(3)
    assert t["z"] > 0 and t["x"] > 0 and t["y"] < 0.05 and t["y_inv"] < 0.05
    res = s.autotune(2)
    assert res is not None
    assert s._lib.ps_is_2d(s._ctx) in (1, 2)
    b = np.random.default_rng(4).standard_normal((256, 256))
    lg, _ = gpu_solve2d(n, (0, 0), (1 / 256, 1 / 256), 2, [b], "plain", solver=s)
    lo, _ = O.solve2d(n, (0, 0), (1 / 256, 1 / 256), 2, [b])
    assert O.rel_l2(lg[0], lo[0]) < 2e-6
    s.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 1e-6)])
def test_2d_4096_specialised_8192(prec, tol):
    """4096^2 unbounded (transform length 8192, the specialised 8192 kernels):
    grad of the reference's unbounded2d bump against its analytic gradient."""
    n = 4096
    x, h = cell_centres(n, -1.0, 1.0)
    X, Y = np.meshgrid(x, x)
    A, B, Ax, Ay = bump(X, Y)
    announce("2d 4096", prec)
    lg, _ = gpu_solve2d((n, n), (0, 0), (h, h), 0, [B], "grad", prec=prec)
    assert O.rel_l2([lg[0], lg[1]], [Ax, Ay]) < tol


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
@pytest.mark.parametrize("op", ["plain", "grad"])
def test_2d_two_residue_y_pass(prec, op):
    """ny = 1024 unbounded: the y transform has length 2048, run as two
    1024-point residues of the unpadded data: on the FP32 two-residue kernel
    (FP32 form) and on the matrix cores (mixed form)."""
    n = (48, 1024)
    dx = (1 / 48, 1 / 512)
    s = P.poisson_solver(P.PS_FP32 if prec == "fp32" else P.PS_MIXED)
    s.setup2d(n, (0, 0), dx, 2)
    assert s.zpass_variant() == ("two-residue" if prec == "fp32" else "matrix-core")
    s.close()
    _compare(n, (0, 0), dx, 2, op, prec, seed=13)


LONG_Y = [
    # (ncell, bc, dx): y transforms of 8192 / 4096 (one or two columns per workgroup)
    ((40, 4096), (0, 0), (1 / 40, 1 / 2048)),
    ((24, 2048), (1, 0), (1 / 24, 1 / 1024)),
    ((30, 4096), (0, 1), (0.05, 1 / 4096)),
    ((16, 4096), (1, 1), (1 / 16, 1 / 4096)),
]


@pytest.mark.parametrize("case", LONG_Y, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("op", ["plain", "grad", "div", "curl"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("t2d", [1, 0])
def test_2d_long_y_transposed_and_in_place(case, op, prec, t2d):
    """Long y pencils with the fused pass reading a transposed copy of the x
    pass output (t2d 1: contiguous pencils, grad / curl transform their scalar
    once) and in place on the x pass layout (t2d 0), against the oracle."""
    n, bc, dx = case
    s = P.poisson_solver(P.PS_FP64 if prec == "fp64" else P.PS_FP32)
    s.set_tuning(t2d=t2d)
    s.setup2d(n, bc, dx, 2)
    assert s.last_status == 0
    assert s._lib.ps_is_2d(s._ctx) == (2 if t2d else 1)
    announce("2d long y", n, bc, op, prec, "t2d", t2d)
    b = _rhs(n, 29, op)
    lg, rg = gpu_solve2d(n, bc, dx, 2, b, op=op, prec=prec, want_rhs=True, solver=s)
    s.close()
    lo, ro = O.solve2d(n, bc, dx, 2, b, op=op)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= TOL[prec]
    for d in range(3):
        if lo[d] is None:
            assert np.isnan(lg[d]).all()
    gr = [rg[d] for d in range(3) if ro[d] is not None]
    rr = [ro[d] for d in range(3) if ro[d] is not None]
    assert O.rel_l2(gr, rr) <= TOL[prec]


def test_2d_transposed_path_repeated_solves_and_ops():
    """One context, one transposed set-up: plain, grad, curl, div and plain
    again in sequence (buffers grow from one to three components between
    solves), each against the oracle."""
    n, bc, dx = (20, 4096), (0, 0), (0.05, 1 / 2048)
    s = P.poisson_solver(P.PS_FP64)
    s.setup2d(n, bc, dx, 0)
    assert s._lib.ps_is_2d(s._ctx) == 2
    for i, op in enumerate(["plain", "grad", "curl", "div", "plain"]):
        b = _rhs(n, 40 + i, op)
        lg, _ = gpu_solve2d(n, bc, dx, 0, b, op=op, prec="fp64", solver=s)
        lo, _ = O.solve2d(n, bc, dx, 0, b, op=op)
        got = [lg[d] for d in range(3) if lo[d] is not None]
        ref = [lo[d] for d in range(3) if lo[d] is not None]
        assert O.rel_l2(got, ref) <= 1e-12, op
    s.close()
