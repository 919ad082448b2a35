"""The potential extension (ps_solve3d_potential): A, grad A and the
Laplacian of A from one pushed scalar in one pipeline, against the oracle
(A: plain solve, grad: the reference's component-wise grad of the scalar
pushed three times, Laplacian: (ik.ik) G r, which the oracle pins to the
closed form -Gaussian(s_eff) for the config-1 source) and against the
component-wise GPU solves it replaces."""
import math

import numpy as np
import pytest

import bench
import poissonsolver_amd as P
from oracle import oracle as O
from tests.gpu_util import PREC, announce, gpu_solve, random_rhs

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}
NAMES = ("A", "dx", "dy", "dz", "lap")


def _potential(prec, n, bc, dx, reg, b, outputs=NAMES, tuning=None, zlaunches=None):
    s = P.poisson_solver(PREC[prec])
    if tuning:
        s.set_tuning(**tuning)
    s.setup3d(n, bc, dx, reg)
    assert s.last_status == 0
    if zlaunches is not None:
        assert s.potential_zlaunches(outputs) == zlaunches
    s.push(b)
    s.solve3d_potential(outputs)
    assert s.last_status == 0
    out = {k: np.full(b.shape, np.nan) for k in NAMES}
    s.pull_potential(*[out[k] for k in NAMES])
    assert s.last_status == 0
    rb = np.full(b.shape, np.nan)
    s.pull(None, None, None, rb)          # the held rhs is left as pushed
    s.close()
    return out, rb


CASES = [((12, 10, 8), (0, 0, 0), (0.1, 0.12, 0.09), 2),
         ((16, 12, 10), (1, 1, 0), (1 / 16, 1 / 12, 0.1), 6),
         ((10, 9, 12), (0, 0, 1), (0.1, 0.1, 1 / 12), 0),
         ((8, 10, 12), (1, 1, 1), (1 / 8, 0.1, 1 / 12), 4),
         ((16, 12, 512), (0, 0, 0), (0.1, 0.1, 1 / 256), 2)]   # Mz = 1024: matrix-core z pass when mixed


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%dx%d-%d%d%d" % (c[0] + c[1]))
@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_potential_vs_oracle_and_componentwise(case, prec):
    n, bc, dx, reg = case
    announce("potential", n, bc, reg, prec)
    b = random_rhs(n, 77, 1)[0]
    out, rb = _potential(prec, n, bc, dx, reg, b)
    tol = TOL[prec]
    lo_a, _ = O.solve(n, bc, dx, reg, [b])
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, reg, [b], op="lap")
    assert O.rel_l2(out["A"], lo_a[0]) <= tol
    assert O.rel_l2([out["dx"], out["dy"], out["dz"]], lo_g) <= tol
    assert O.rel_l2(out["lap"], lo_l[0]) <= tol
    assert np.array_equal(rb, b if prec == "fp64" else b.astype(np.float32).astype(np.float64))
    # the component-wise solves it replaces
    ga, _ = gpu_solve(n, bc, dx, reg, [b], prec=prec)
    gg, _ = gpu_solve(n, bc, dx, reg, [b, b, b], op="grad", prec=prec)
    assert O.rel_l2(out["A"], ga[0]) <= tol
    assert O.rel_l2([out["dx"], out["dy"], out["dz"]], gg) <= tol


def test_potential_subset_and_errors():
    n, bc, dx, reg = CASES[0]
    b = random_rhs(n, 78, 1)[0]
    out, _ = _potential("fp64", n, bc, dx, reg, b, outputs=("dy", "lap"))
    for k in ("A", "dx", "dz"):
        assert np.isnan(out[k]).all(), k
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    assert O.rel_l2(out["dy"], lo_g[1]) <= 1e-12
    s = P.poisson_solver()
    assert s.potential_zlaunches(("A",)) == -P.PS_ESTATE       # not set up
    s.setup3d(n, bc, dx, reg)
    # FP64: one z launch per distinct z multiplier (A with d/dx and d/dy, d/dz, Laplacian)
    assert s.potential_zlaunches(("A",)) == 1
    assert s.potential_zlaunches(("dx", "dy", "dz")) == 2
    assert s.potential_zlaunches(NAMES) == 3
    s.solve3d_potential(("A",))
    assert s.last_status == P.PS_ESTATE          # nothing pushed
    s.push(None, b, None)
    s.solve3d_potential(("A",))
    assert s.last_status == P.PS_ESTATE          # the scalar is rhs X
    s.close()


def test_potential_laplacian_closed_form():
    """Gaussian source, reg 2 (config 1): the Laplacian of the regularised
    potential is minus the Gaussian of width s_eff = sqrt(s^2 + (2h)^2)."""
    n = 64
    announce("potential closed form", n)
    b, h = bench.gaussian_source(n)
    out, _ = _potential("fp64", (n, n, n), (0, 0, 0), (h, h, h), 2, b, outputs=("lap",))
    se = math.sqrt(0.1 ** 2 + (2 * h) ** 2)
    x = -1.0 + (np.arange(n) + 0.5) * h
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    g = np.exp(-r2 / (2 * se * se)) / ((2 * np.pi) ** 1.5 * se ** 3)
    assert O.rel_l2(out["lap"], -g) <= 1e-10


SUBSETS = [("dx",), ("dy",), ("dx", "dy"), ("A", "dy"), ("dx", "dz", "lap"), NAMES]


@pytest.mark.parametrize("zblk", [-1, 0, 3, 5], ids=lambda z: "zblk%d" % z)
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_potential_subsets_in_every_spectrum_layout(zblk, prec):
    """d/dx A and d/dy A are derived from the z-pass spectrum of A in the
    y-pass layout of the spectrum buffer: every output subset (with and
    without A itself) in the three layouts (ky-major, z-major, z blocks)."""
    n, bc, dx, reg = (12, 10, 40), (0, 0, 0), (0.1, 0.12, 0.03), 2
    b = random_rhs(n, 79, 1)[0]
    lo_a, _ = O.solve(n, bc, dx, reg, [b])
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, reg, [b], op="lap")
    ref = {"A": lo_a[0], "dx": lo_g[0], "dy": lo_g[1], "dz": lo_g[2], "lap": lo_l[0]}
    for outs in SUBSETS:
        announce("potential subset", n, prec, zblk, outs)
        out, _ = _potential(prec, n, bc, dx, reg, b, outputs=outs, tuning={"zblk": zblk})
        for k in NAMES:
            if k in outs:
                assert O.rel_l2(out[k], ref[k]) <= TOL[prec], (outs, k)
            else:
                assert np.isnan(out[k]).all(), (outs, k)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_potential_long_z(prec):
    """z transforms of 2048 (nz = 1024 unbounded; the matrix-core kernel of
    two 1024-point residues when mixed), 32-plane z blocks."""
    n, bc, dx, reg = (8, 6, 1024), (0, 0, 0), (0.1, 0.12, 1 / 512), 4
    announce("potential long z", n, prec)
    b = random_rhs(n, 80, 1)[0]
    out, _ = _potential(prec, n, bc, dx, reg, b)
    lo_a, _ = O.solve(n, bc, dx, reg, [b])
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, reg, [b], op="lap")
    assert O.rel_l2(out["A"], lo_a[0]) <= 2e-6
    assert O.rel_l2([out["dx"], out["dy"], out["dz"]], lo_g) <= 2e-6
    assert O.rel_l2(out["lap"], lo_l[0]) <= 2e-6


FAN_SUBSETS = [("A", "dz"), ("A", "lap"), ("dz", "lap"), ("dx", "dz"), ("dy", "lap"), ("A", "dx", "dy", "dz"),
               NAMES]
FAN_CASES = [((18, 10, 256), (0, 0, 0), (0.1, 0.12, 1 / 128), 2),
             ((16, 12, 256), (1, 1, 0), (1 / 16, 1 / 12, 1 / 128), 6),
             ((10, 6, 1024), (0, 0, 0), (0.1, 0.12, 1 / 512), 4),
             ((12, 7, 1024), (1, 1, 0), (1 / 12, 1 / 7, 1 / 512), 10)]


@pytest.mark.parametrize("case", FAN_CASES, ids=lambda c: "%dx%dx%d-%d%d%d" % (c[0] + c[1]))
@pytest.mark.parametrize("zblk", [-1, 0, 3], ids=lambda z: "zblk%d" % z)
def test_potential_fanout(case, zblk):
    """z transforms of 512 and 2048 (nz = 256, 1024 unbounded), mixed form: the z outputs
    (A, d/dz A, Laplacian) from one forward transform of the scalar (tuning
    zfan 1: every two- and three-output set, ragged column tiles, odd and
    even slab counts) against the oracle (rel-L2 <= 1e-6) and against one z
    launch per output (zfan 0; the same arithmetic, rel-L2 <= 1e-7)."""
    n, bc, dx, reg = case
    b = random_rhs(n, 81, 1)[0]
    lo_a, _ = O.solve(n, bc, dx, reg, [b])
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, reg, [b], op="lap")
    ref = {"A": lo_a[0], "dx": lo_g[0], "dy": lo_g[1], "dz": lo_g[2], "lap": lo_l[0]}
    for outs in FAN_SUBSETS:
        announce("potential fan-out", n, bc, zblk, outs)
        nzq = int(bool({"A", "dx", "dy"} & set(outs))) + int("dz" in outs) + int("lap" in outs)
        f1, _ = _potential("mixed", n, bc, dx, reg, b, outputs=outs, tuning={"zblk": zblk, "zfan": 1}, zlaunches=1)
        f0, _ = _potential("mixed", n, bc, dx, reg, b, outputs=outs, tuning={"zblk": zblk, "zfan": 0},
                           zlaunches=nzq)
        for k in NAMES:
            if k in outs:
                assert O.rel_l2(f1[k], ref[k]) <= TOL["mixed"], (outs, k)
                assert O.rel_l2(f1[k], f0[k]) <= 1e-7, (outs, k)
            else:
                assert np.isnan(f1[k]).all(), (outs, k)


YFAN_SUBSETS = [("A", "dx", "dy"), ("dx", "dy"), ("A", "dy"), ("A", "dx", "dz"), ("dx",), NAMES]
# (n, bc, dx, reg, x length specialised, y length specialised)
YFAN_CASES = [((12, 32, 10), (0, 0, 0), (0.1, 1 / 32, 0.1), 2, False, True),
              ((10, 64, 12), (1, 1, 0), (0.1, 1 / 64, 0.08), 6, False, True),
              ((32, 32, 10), (0, 0, 0), (1 / 32, 1 / 32, 0.1), 2, True, True),
              ((64, 12, 10), (1, 1, 0), (1 / 64, 0.08, 0.1), 4, True, False)]


def _ylaunch_plan(outs, zfan, xstat, ystat, prec):
    """(y^-1 launches, y^-1 outputs) of the potential form, as the library plans them."""
    xdx = bool(zfan and xstat and "dx" in outs)
    ents = [q for q in ("A", "dx", "dy") if q in outs and not (q == "dx" and xdx and "A" in outs)]
    fan = len(ents) >= 3 and prec != "fp64" and ystat and zfan
    rest = int("dz" in outs) + int("lap" in outs)
    return (1 if fan else len(ents)) + rest, len(ents) + rest


@pytest.mark.parametrize("case", YFAN_CASES, ids=lambda c: "%dx%dx%d-%d%d%d" % (c[0] + c[1]))
@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_potential_y_fanout(case, prec):
    """Specialised x and y lengths (64): the y^-1 passes of A, dA/dy (and
    dA/dx) from one load of A's z output, and dA/dx from A's y^-1 output in
    the x^-1 pass (times i kx), tuning zfan 1, against the oracle and against
    one y^-1 launch per output (zfan 0); ps_potential_ylaunches and
    ps_potential_youtputs assert which form ran."""
    n, bc, dx, reg, xstat, ystat = case
    b = random_rhs(n, 83, 1)[0]
    lo_a, _ = O.solve(n, bc, dx, reg, [b])
    lo_g, _ = O.solve(n, bc, dx, reg, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, reg, [b], op="lap")
    ref = {"A": lo_a[0], "dx": lo_g[0], "dy": lo_g[1], "dz": lo_g[2], "lap": lo_l[0]}
    for outs in YFAN_SUBSETS:
        announce("potential y fan-out", n, bc, prec, outs)
        res = {}
        for zfan in (1, 0):
            s = P.poisson_solver(PREC[prec])
            s.set_tuning(zfan=zfan)
            s.setup3d(n, bc, dx, reg)
            want_l, want_o = _ylaunch_plan(outs, zfan, xstat, ystat, prec)
            assert s.potential_ylaunches(outs) == want_l, (outs, zfan)
            assert s.potential_youtputs(outs) == want_o, (outs, zfan)
            s.push(b)
            s.solve3d_potential(outs)
            assert s.last_status == 0
            out = {k: np.full(b.shape, np.nan) for k in NAMES}
            s.pull_potential(*[out[k] for k in NAMES])
            s.close()
            res[zfan] = out
        for k in NAMES:
            if k in outs:
                assert O.rel_l2(res[1][k], ref[k]) <= TOL[prec], (outs, k)
                assert O.rel_l2(res[1][k], res[0][k]) <= 5e-7, (outs, k)
            else:
                assert np.isnan(res[1][k]).all(), (outs, k)
