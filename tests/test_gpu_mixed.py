"""Mixed-precision form (PS_MIXED) vs the oracle: the z pass of a 256-, 512-
or 1024-cell unbounded z direction (transform length 512 = 32 x 16, 1024 =
32 x 32, or 2048 as two 1024-point residues) runs on the matrix cores as
FP16 hi/lo DFT-32 and DFT-16 products with FP32 accumulation; every other
pass is FP32.
Bar: rel-L2 <= 1e-6 against the FP64 oracle (the tolerance of the
reduced-precision form).  x and y stay small so the oracle finishes in
seconds."""
import numpy as np
import pytest

import poissonsolver_amd as P
from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, gpu_solve2d, random_rhs

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6

CASES = [
    ((8, 6, 512), (0, 0, 0), (0.1, 0.12, 2 / 512)),
    ((12, 10, 512), (0, 0, 0), (0.05, 0.06, 1 / 256)),
    ((16, 12, 512), (1, 1, 0), (1 / 16, 1 / 12, 2 / 512)),
    ((9, 7, 512), (0, 0, 0), (0.1, 0.1, 0.004)),
]


def _variant(n, bc, dx, reg):
    s = P.poisson_solver(P.PS_MIXED)
    s.setup3d(n, bc, dx, reg)
    v = s.zpass_variant()
    s.close()
    return v


def _compare(n, bc, dx, reg, op, b, want_rhs=False, tol=MIXED_TOL):
    lo, ro = O.solve(n, bc, dx, reg, b, op=op, reproject=False)
    lg, rg = gpu_solve(n, bc, dx, reg, b, op=op, prec="mixed", want_rhs=want_rhs)
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    err = O.rel_l2(got, ref)
    assert err <= tol, ("lhs", err)
    if want_rhs:
        gr = [rg[d] for d in range(3) if ro[d] is not None]
        rr = [ro[d] for d in range(3) if ro[d] is not None]
        err_r = O.rel_l2(gr, rr)
        assert err_r <= tol, ("rhs", err_r)
    return err


# matrix-core kernel forms (PS_TUNE_MFMA) of transforms of 512 and 1024: 1
# the register-pencil kernel, 2 one tile per workgroup
FORMS = ["1", "2"]
# of transforms of 2048: 1 the register-pencil kernel, 2 eight-pencil and
# 3 four-pencil workgroups
FORMS2048 = ["1", "2", "3"]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("reg", [0, 2, 6])
@pytest.mark.parametrize("form", FORMS)
def test_mixed_plain(case, reg, form, monkeypatch):
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = case
    announce("mixed plain", n, bc, reg, "form", form)
    assert _variant(n, bc, dx, reg) == "matrix-core"
    b = random_rhs(n, 23)
    _compare(n, bc, dx, reg, "plain", b)


@pytest.mark.parametrize("case", CASES[:3], ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("form", FORMS)
def test_mixed_grad_and_filtered_rhs(case, form, monkeypatch):
    # grad: one-component z operator i k_d Gk per component; the pulled rhs
    # with reg > 0 runs the filter pipeline (zeta / N) through the same kernel
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = case
    announce("mixed grad + filtered rhs", n, bc, "form", form)
    b = random_rhs(n, 29)
    _compare(n, bc, dx, 2, "grad", b, want_rhs=True)


@pytest.mark.parametrize("form", FORMS)
def test_mixed_dynamic_range(form, monkeypatch):
    # pencils whose magnitudes span 2^-60 .. 2^60: every stage scales its fp16
    # operands per pencil or per DFT column
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = (16, 8, 512), (0, 0, 0), (0.1, 0.1, 2 / 512)
    announce("mixed dynamic range", n)
    rng = np.random.default_rng(31)
    amp = np.exp2(np.linspace(-60, 60, n[0]))[None, None, :] * np.exp2(np.linspace(-20, 20, n[1]))[None, :, None]
    b = [rng.standard_normal((n[2], n[1], n[0])) * amp]
    _compare(n, bc, dx, 0, "plain", b)
    # a source at 2^+-30 times that (still inside the FP32 range) gives the
    # solution scaled by the same power of two
    b2 = rng.standard_normal((n[2], n[1], n[0])) * np.exp2(np.linspace(-30, 30, n[0]))[None, None, :]
    lg1, _ = gpu_solve(n, bc, dx, 0, [b2 * 2.0 ** 40], prec="mixed")
    lg2, _ = gpu_solve(n, bc, dx, 0, [b2 * 2.0 ** -40], prec="mixed")
    assert np.isfinite(lg1[0]).all() and np.isfinite(lg2[0]).all()
    assert O.rel_l2([lg1[0] * 2.0 ** -80], [lg2[0]]) <= 1e-6


def test_mixed_zero_source():
    n, bc, dx = (8, 8, 512), (0, 0, 0), (0.1, 0.1, 2 / 512)
    announce("mixed zero source", n)
    lg, _ = gpu_solve(n, bc, dx, 2, [np.zeros((n[2], n[1], n[0]))], prec="mixed")
    assert np.all(lg[0] == 0.0)


def test_mixed_other_lengths_use_fp32_kernels():
    n, bc, dx = (16, 12, 20), (0, 0, 0), (0.1, 0.1, 0.1)
    announce("mixed fallback", n)
    assert _variant(n, bc, dx, 2) != "matrix-core"
    b = random_rhs(n, 37)
    _compare(n, bc, dx, 2, "curl", b, tol=2e-6)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("op", ["curl", "div"])
@pytest.mark.parametrize("form", FORMS)
def test_mixed_three_component_operators(case, op, form, monkeypatch):
    # curl and div: the fused three-component matrix-core z kernels, the
    # register-pencil one (1: one wavefront per pencil column and its three
    # components) and the 4-pencil one-tile one (2)
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = case
    announce("mixed", op, n, bc, "form", form)
    b = random_rhs(n, 41)
    _compare(n, bc, dx, 2, op, b)


@pytest.mark.parametrize("zblk", ["-1", "0", "3", "5"], ids=lambda z: "zblk" + z)
@pytest.mark.parametrize("op", ["curl", "div"])
def test_mixed_three_component_every_layout(zblk, op, monkeypatch):
    """The register-pencil three-component kernel in each z layout of the
    spectrum buffer, on components of magnitudes 2^-30, 1 and 2^30 (each
    component's input scale enters the k-space combination)."""
    monkeypatch.setenv("PS_TUNE_ZBLK", zblk)
    n, bc, dx = (10, 8, 512), (0, 0, 0), (0.1, 0.12, 2 / 512)
    announce("mixed three-component layout", op, n, zblk)
    rng = np.random.default_rng(53)
    b = [rng.standard_normal((n[2], n[1], n[0])) * 2.0 ** sc for sc in (-30, 0, 30)]
    _compare(n, bc, dx, 2, op, b)


def test_mixed_curl_with_reprojection_and_filtered_rhs():
    # the reprojected rhs filter (three-component, reg > 0) runs the FP32
    # kernels; the curl itself runs on the matrix cores
    n, bc, dx = (8, 6, 512), (0, 0, 0), (0.1, 0.12, 2 / 512)
    announce("mixed curl + reprojection", n)
    b = random_rhs(n, 43)
    lo, ro = O.solve(n, bc, dx, 2, b, op="curl", reproject=True)
    lg, rg = gpu_solve(n, bc, dx, 2, b, op="curl", reproject=True, prec="mixed", want_rhs=True)
    assert O.rel_l2(lg, lo) <= MIXED_TOL
    assert O.rel_l2(rg, ro) <= 2e-6


CASES2048 = [
    ((8, 6, 1024), (0, 0, 0), (0.1, 0.12, 1 / 512)),
    ((12, 10, 1024), (1, 1, 0), (1 / 12, 0.1, 1 / 512)),
]


@pytest.mark.parametrize("case", CASES2048, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("reg", [0, 2, 10])
@pytest.mark.parametrize("form", FORMS2048)
def test_mixed_2048_plain_grad_filtered_rhs(case, reg, form, monkeypatch):
    """nz = 1024 unbounded: z transforms of 2048 as two 1024-point residues
    on the matrix cores (plain, component-wise grad, filtered rhs), each
    kernel form; rel-L2 <= 1e-6 against the oracle."""
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = case
    announce("mixed 2048", n, bc, reg, "form", form)
    assert _variant(n, bc, dx, reg) == "matrix-core"
    b = random_rhs(n, 61 + reg)
    _compare(n, bc, dx, reg, "plain", [b[0]], want_rhs=True)
    _compare(n, bc, dx, reg, "grad", b, want_rhs=True)


@pytest.mark.parametrize("form", FORMS2048)
def test_mixed_2048_potential_and_range(form, monkeypatch):
    """The potential extension's five quantities on each 2048-point
    matrix-core kernel form, and a source spanning 2^-40 .. 2^40 in
    magnitude; rel-L2 <= 1e-6 against the oracle."""
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = CASES2048[0]
    announce("mixed 2048 potential", n, "form", form)
    rng = np.random.default_rng(67)
    b = rng.standard_normal((n[2], n[1], n[0])) * np.exp2(rng.uniform(-40, 40, (n[2], 1, 1)))
    s = P.poisson_solver(P.PS_MIXED)
    s.setup3d(n, bc, dx, 2)
    assert s.zpass_variant() == "matrix-core"
    s.push(b)
    s.solve3d_potential(("A", "dx", "dy", "dz", "lap"))
    got = [np.empty_like(b) for _ in range(5)]
    s.pull_potential(*got)
    assert s.last_status == 0
    s.close()
    lo_a, _ = O.solve(n, bc, dx, 2, [b])
    lo_g, _ = O.solve(n, bc, dx, 2, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, 2, [b], op="lap")
    assert O.rel_l2(got[0], lo_a[0]) <= MIXED_TOL
    assert O.rel_l2(got[1:4], lo_g) <= MIXED_TOL
    assert O.rel_l2(got[4], lo_l[0]) <= MIXED_TOL


CASES512 = [
    ((8, 6, 256), (0, 0, 0), (0.1, 0.12, 1 / 128)),
    ((12, 10, 256), (1, 1, 0), (1 / 12, 0.1, 1 / 128)),
    ((9, 7, 256), (0, 0, 0), (0.1, 0.1, 0.008)),
]


@pytest.mark.parametrize("case", CASES512, ids=lambda c: "%s-%s" % (c[1], "x".join(map(str, c[0]))))
@pytest.mark.parametrize("reg", [0, 2, 10])
@pytest.mark.parametrize("form", FORMS)
def test_mixed_512_plain_grad_filtered_rhs(case, reg, form, monkeypatch):
    """nz = 256 unbounded: z transforms of 512 = 32 x 16 on the matrix cores
    (plain, component-wise grad, filtered rhs), on the register-pencil (1) and
    the one-tile (2) kernel; rel-L2 <= 1e-6 against the oracle."""
    monkeypatch.setenv("PS_TUNE_MFMA", form)
    n, bc, dx = case
    announce("mixed 512", n, bc, reg, "form", form)
    assert _variant(n, bc, dx, reg) == "matrix-core"
    b = random_rhs(n, 71 + reg)
    _compare(n, bc, dx, reg, "plain", [b[0]], want_rhs=True)
    _compare(n, bc, dx, reg, "grad", b, want_rhs=True)


@pytest.mark.parametrize("zblk", [-1, 0, 3, 4, 5], ids=lambda z: "zblk%d" % z)
def test_mixed_512_every_layout_and_range(zblk):
    """The 512-point kernel in every z layout of the spectrum buffer, on a
    source spanning 2^-40 .. 2^40 in magnitude, with the potential form's
    quantities (plain, d/dz, Laplacian; d/dx and d/dy from A)."""
    n, bc, dx = (10, 8, 256), (0, 0, 0), (0.1, 0.12, 1 / 128)
    announce("mixed 512 layout", n, zblk)
    rng = np.random.default_rng(73)
    b = rng.standard_normal((n[2], n[1], n[0])) * np.exp2(rng.uniform(-40, 40, (n[2], 1, 1)))
    s = P.poisson_solver(P.PS_MIXED)
    s.set_tuning(zblk=zblk)
    s.setup3d(n, bc, dx, 2)
    assert s.zpass_variant() == "matrix-core"
    s.push(b)
    s.solve3d_potential(("A", "dx", "dy", "dz", "lap"))
    got = [np.empty_like(b) for _ in range(5)]
    s.pull_potential(*got)
    assert s.last_status == 0
    s.close()
    lo_a, _ = O.solve(n, bc, dx, 2, [b])
    lo_g, _ = O.solve(n, bc, dx, 2, [b, b, b], op="grad")
    lo_l, _ = O.solve(n, bc, dx, 2, [b], op="lap")
    assert O.rel_l2(got[0], lo_a[0]) <= MIXED_TOL
    assert O.rel_l2(got[1:4], lo_g) <= MIXED_TOL
    assert O.rel_l2(got[4], lo_l[0]) <= MIXED_TOL


def test_mixed_512_2d():
    """2D: y transforms of 512 (ny = 256 unbounded) on the 512-point kernel."""
    n, bc, dx = (24, 256), (0, 0), (0.05, 1 / 128)
    announce("mixed 512 2D", n)
    b = np.random.default_rng(79).standard_normal((n[1], n[0]))
    s = P.poisson_solver(P.PS_MIXED)
    s.setup2d(n, bc, dx, 2)
    assert s.zpass_variant() == "matrix-core"
    lg, _ = gpu_solve2d(n, bc, dx, 2, [b], op="grad", prec="mixed", solver=s)
    s.close()
    lo, _ = O.solve2d(n, bc, dx, 2, [b], op="grad")
    assert O.rel_l2([lg[0], lg[1]], [lo[0], lo[1]]) <= MIXED_TOL
