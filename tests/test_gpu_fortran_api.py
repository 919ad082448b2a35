"""The reference's Fortran interface over the C ABI: one-array push/pull with
ghost offsets (poisson_solver_push.f:8, poisson_solver_pull.f:8) and
smooth3d (poisson_solver_smooth3d.f:8).  Strided push/pull must give the
planar path's results bit for bit and leave ghost cells untouched;
smooth3d is compared with the oracle's restatement (FP64 rel-L2 <= 1e-12,
FP32 <= 2e-6, mixed <= 1e-6)."""
import numpy as np
import pytest

import poissonsolver_amd as P
from poissonsolver_amd import _capi
from oracle import oracle as O
from tests.gpu_util import PREC, announce, gpu_solve, gpu_solve2d

pytestmark = pytest.mark.gpu

SENTINEL = 777.25


def _block(shape_zyx, ncomp, g, fill=SENTINEL):
    nz, ny, nx = shape_zyx
    return np.full((nz + 2 * g, ny + 2 * g, nx + 2 * g, ncomp), fill)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("ncomp,op", [(1, "plain"), (2, "plain"), (3, "curl"), (3, "grad")])
def test_strided_push_pull_matches_planar(prec, ncomp, op):
    n = (12, 10, 9)
    bc, dx, reg, g = (0, 0, 0), (0.1, 0.12, 0.09), 2, 2
    announce("strided", prec, ncomp, op)
    rng = np.random.default_rng(3)
    b = [rng.standard_normal((n[2], n[1], n[0])) for _ in range(ncomp)]
    ref_l, ref_r = gpu_solve(n, bc, dx, reg, b, op=op, prec=prec, want_rhs=True)
    blk = _block((n[2], n[1], n[0]), ncomp, g)
    for c in range(ncomp):
        blk[g:-g, g:-g, g:-g, c] = b[c]
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, reg)
    s.set_return_grad(op == "grad")
    s.set_return_curl(op == "curl")
    s.push_strided(blk, (g, g, g))
    assert s.last_status == 0
    s.solve3d()
    assert s.last_status == 0
    nl = 3
    lb = _block((n[2], n[1], n[0]), nl, g)
    rb = _block((n[2], n[1], n[0]), ncomp, g)
    s.pull_strided(lb, rb, (g, g, g))
    assert s.last_status == 0
    s.close()
    for c in range(nl):
        if np.isnan(ref_l[c]).all():
            assert (lb[..., c] == SENTINEL).all()          # component not held: untouched
        else:
            assert np.array_equal(lb[g:-g, g:-g, g:-g, c], ref_l[c])
    for c in range(ncomp):
        assert np.array_equal(rb[g:-g, g:-g, g:-g, c], ref_r[c])
    for arr in (lb, rb):
        inner = np.zeros(arr.shape[:3], bool)
        inner[g:-g, g:-g, g:-g] = True
        assert (arr[~inner] == SENTINEL).all()             # ghost cells untouched


def test_strided_asymmetric_offsets_2d():
    """A 2D context: block (ny', nx', ncomp) with different ghost widths per side."""
    n = (16, 12)
    dx = (0.1, 0.08)
    rng = np.random.default_rng(8)
    b = rng.standard_normal((n[1], n[0]))
    ref, _ = gpu_solve2d(n, (0, 0), dx, 0, [b], "grad")
    blk = np.full((n[1] + 3, n[0] + 4, 1), SENTINEL)
    blk[1:1 + n[1], 3:3 + n[0], 0] = b
    s = P.poisson_solver()
    s.setup2d(n, (0, 0), dx, 0)
    s.set_return_grad(True)
    s.push_strided(blk, (3, 1))
    s.solve2d()
    out = np.full((n[1] + 3, n[0] + 4, 2), SENTINEL)
    s.pull_strided(out, None, (3, 1))
    assert s.last_status == 0
    for c in range(2):
        assert np.array_equal(out[1:1 + n[1], 3:3 + n[0], c], ref[c])
    mask = np.ones(out.shape[:2], bool)
    mask[1:1 + n[1], 3:3 + n[0]] = False
    assert (out[mask] == SENTINEL).all()
    s.close()


def test_strided_argument_errors():
    s = P.poisson_solver()
    s.setup3d((8, 8, 8), (0, 0, 0), (0.1,) * 3)
    s.push_strided(np.zeros((8, 8, 8, 4)))
    assert s.last_status == _capi.PS_EINVAL
    s.close()


CASES = [
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125)),
    ((16, 12, 20), (1, 1, 0), (1 / 16, 1 / 12, 0.1)),
    ((14, 10, 7), (0, 0, 1), (0.11, 0.09, 1 / 7)),
    ((9, 7, 5), (1, 1, 1), (1 / 9, 1 / 7, 0.2)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s" % (c[1],))
@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_smooth3d_vs_oracle(case, prec):
    n, bc, dx = case
    announce("smooth3d", n, bc, prec)
    rng = np.random.default_rng(4)
    b = [rng.standard_normal((n[2], n[1], n[0])) for _ in range(2)]
    h = 1.5 * max(dx)
    o = O.OracleSolver()
    o.setup3d(n, bc, dx, 0)
    o.push(b[0], None, b[1])
    o.smooth3d(h)
    want = [o.rhs[d][:, :, :n[0]].real for d in (0, 2)]
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, 0)
    s.push(b[0], None, b[1])
    s.smooth3d(h)
    assert s.last_status == 0
    got = [np.full((n[2], n[1], n[0]), np.nan) for _ in range(3)]
    s.pull(None, None, None, *got)
    assert np.isnan(got[1]).all()
    tol = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}[prec]
    assert O.rel_l2([got[0], got[2]], want) < tol
    # the next solve uses the smoothed field, x-extended as the routine
    # leaves it in its x pencils
    s.solve3d()
    lg = [np.full((n[2], n[1], n[0]), np.nan) for _ in range(3)]
    s.pull(*lg)
    o.solve3d()
    lo = [o.lhs[d][:, :, :n[0]].real for d in (0, 2)]
    assert O.rel_l2([lg[0], lg[2]], lo) < tol
    s.close()


def test_smooth3d_matrix_core_and_identity():
    """nz = 512 unbounded in the mixed form: the z pass of smooth3d runs on the
    matrix cores; h = 0 is the identity."""
    n = (8, 8, 512)
    dx = (0.1, 0.1, 1 / 256)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((n[2], n[1], n[0]))
    o = O.OracleSolver()
    o.setup3d(n, (0, 0, 0), dx, 0)
    o.push(b)
    o.smooth3d(0.02)
    s = P.poisson_solver(P.PS_MIXED)
    s.setup3d(n, (0, 0, 0), dx, 0)
    assert s.zpass_variant() == "matrix-core"
    s.push(b)
    s.smooth3d(0.02)
    got = np.empty_like(b)
    s.pull(None, None, None, got)
    assert O.rel_l2(got, o.rhs[0][:, :, :n[0]].real) < 1e-6
    s.close()
    s = P.poisson_solver()
    s.setup3d(n, (0, 0, 0), dx, 0)
    s.push(b)
    s.smooth3d(0.0)
    s.pull(None, None, None, got)
    assert O.rel_l2(got, b) < 1e-14
    s.smooth3d(-1.0)
    assert s.last_status == _capi.PS_EINVAL
    s.close()
    s = P.poisson_solver()
    s.setup2d((8, 8), (0, 0), (0.1, 0.1), 0)
    s.smooth3d(0.1)
    assert s.last_status == _capi.PS_ESTATE
    s.close()
