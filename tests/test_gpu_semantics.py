"""Call-sequence semantics of the drop-in against the oracle's restatement of
the reference object's state (rhs/lhs arrays that persist between calls).

* Repeat solves: a solve with regularisation (or reprojection of three
  components) leaves the filtered field in the reference's rhs arrays, z and y
  truncated but x extended over the doubled rows (solve3d.cpp:403-629); the
  next solve maps and transforms that field again (map.cpp:50-60).
* Host transfers are ordered after an asynchronous solve on the context's
  stream.
* A device-memory limit turns allocations beyond it into PS_ENOMEM and leaves
  the context usable.
* NaN / Inf sources give non-finite solutions on every precision form (no
  finite value is invented).
"""
import ctypes

import numpy as np
import pytest

import poissonsolver_amd as P
from oracle import oracle as O
from poissonsolver_amd import _capi
from tests.gpu_util import PREC, DeviceBuffer, announce, random_rhs

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 2e-6, "mixed": 1e-6}


def _ops(s, op, reproject):
    s.set_return_grad(op == "grad")
    s.set_return_div(op == "div")
    s.set_return_curl(op == "curl")
    s.set_reprojection(reproject)


def _run_sequence(n, bc, dx, reg, steps, prec="fp64", dim=3, drop_imag=False):
    """Run `steps` on the GPU solver and on the OracleSolver side by side.
    A step is ("push", [arrays or None]), ("solve", op, reproject),
    ("pull",) or ("smooth", h).  Returns the list of (got, want) pairs of
    every pull and of the lhs/rhs after the last step.  drop_imag: after each
    oracle solve keep only the real part of its rhs arrays."""
    s = P.poisson_solver(PREC[prec])
    o = O.OracleSolver()
    if dim == 3:
        s.setup3d(n, bc, dx, reg)
        o.setup3d(n, bc, dx, reg)
        shape = (n[2], n[1], n[0])
    else:
        s.setup2d(n, bc, dx, reg)
        o.setup2d(n, bc, dx, reg)
        shape = (n[1], n[0])
    assert s.last_status == 0
    pairs = []

    def pull():
        lhs = [np.full(shape, np.nan) for _ in range(3)]
        rhs = [np.full(shape, np.nan) for _ in range(3)]
        s.pull(*lhs, *rhs)
        assert s.last_status == 0
        for d in range(3):
            if o.lhs[d] is not None:
                pairs.append((lhs[d], o.lhs[d][..., :n[0]].real.reshape(shape)))
            if o.rhs[d] is not None:
                pairs.append((rhs[d], o.rhs[d][..., :n[0]].real.reshape(shape)))

    for st in steps:
        if st[0] == "push":
            arrs = list(st[1]) + [None] * (3 - len(st[1]))
            s.push(*arrs)
            o.push(*arrs)
            assert s.last_status == 0
        elif st[0] == "solve":
            _ops(s, st[1], st[2])
            _ops(o, st[1], st[2])
            if dim == 3:
                s.solve3d()
                o.solve3d()
            else:
                s.solve2d()
                o.solve2d()
            assert s.last_status == 0
            if drop_imag:
                o.rhs = [None if r is None else r.real.astype(np.complex128) for r in o.rhs]
        elif st[0] == "smooth":
            s.smooth3d(st[1])
            o.smooth3d(st[1])
            assert s.last_status == 0
        elif st[0] == "pull":
            pull()
    pull()
    s.close()
    return pairs


def _check(pairs, tol):
    """rel-L2 per output; an output that is zero to round-off (the divergence
    of a reprojected field) is measured against the largest output norm."""
    scale = max(float(np.sqrt(np.sum(w ** 2))) for _, w in pairs)
    for k, (got, want) in enumerate(pairs):
        wn = float(np.sqrt(np.sum(want ** 2)))
        if wn < 1e-8 * scale:
            err = float(np.sqrt(np.sum((got - want) ** 2))) / scale
        else:
            err = O.rel_l2(got, want)
        assert err <= tol, (k, err)


CASES3 = [((10, 8, 12), (0, 0, 0), (0.1, 0.12, 0.09)),
          ((12, 10, 8), (1, 1, 0), (1 / 12, 0.1, 0.1)),
          ((9, 8, 10), (0, 0, 1), (0.1, 0.1, 0.1)),
          ((8, 9, 10), (1, 1, 1), (1 / 8, 1 / 9, 0.1))]


@pytest.mark.parametrize("case", CASES3, ids=lambda c: "%d%d%d" % c[1])
@pytest.mark.parametrize("reg", [2, 6])
def test_repeat_solve_with_filter_fp64(case, reg):
    """push -> solve -> solve ... (plain, grad, div with reprojection, curl
    with reprojection): every solve starts from the field the previous one
    left held."""
    n, bc, dx = case
    announce("repeat solve", n, bc, reg)
    b = random_rhs(n, 31 + reg)
    steps = [("push", b), ("solve", "plain", False), ("pull",), ("solve", "plain", False),
             ("solve", "grad", False), ("pull",), ("solve", "div", True), ("solve", "curl", True)]
    _check(_run_sequence(n, bc, dx, reg, steps), 1e-12)


@pytest.mark.parametrize("case", CASES3[:2], ids=lambda c: "%d%d%d" % c[1])
def test_repeat_solve_partial_push_fp64(case):
    """After a filtered solve, pushing one component again replaces only that
    component; the others keep their filtered, x-extended state."""
    n, bc, dx = case
    announce("partial push", n, bc)
    b = random_rhs(n, 41)
    c = random_rhs(n, 42, 1)
    steps = [("push", b), ("solve", "plain", True), ("push", [None, c[0], None]), ("solve", "plain", False),
             ("push", [c[0], None, None]), ("solve", "curl", False)]
    _check(_run_sequence(n, bc, dx, 2, steps), 1e-12)


def test_repeat_solve_reprojection_reg0_plain_fp64():
    """reg 0 with reprojection: the reference's rhs then also carries an
    imaginary part (its reprojection uses the negative Nyquist wavenumber);
    the plain operator and an unfiltered rhs do not see it, so the sequence
    matches the oracle exactly."""
    n, bc, dx = CASES3[0]
    b = random_rhs(n, 51)
    steps = [("push", b), ("solve", "plain", True), ("solve", "plain", False), ("solve", "plain", False)]
    _check(_run_sequence(n, bc, dx, 0, steps), 1e-12)


@pytest.mark.parametrize("case", CASES3, ids=lambda c: "%d%d%d" % c[1])
def test_repeat_solve_held_rhs_is_real_part_fp64(case):
    """The held rhs is the real part of the reference's x pencils: a
    derivative operator after a reg-0 reprojected solve matches the oracle
    whose rhs arrays keep only their real part."""
    n, bc, dx = case
    b = random_rhs(n, 52)
    steps = [("push", b), ("solve", "curl", True), ("solve", "grad", False), ("solve", "curl", True)]
    _check(_run_sequence(n, bc, dx, 0, steps, drop_imag=True), 1e-12)


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_repeat_solve_single_precision(prec):
    n, bc, dx = (16, 12, 512), (0, 0, 0), (0.1, 0.1, 1 / 256)   # nz = 512: matrix-core z pass when mixed
    announce("repeat solve", prec)
    b = random_rhs(n, 61, 1)
    steps = [("push", b), ("solve", "plain", False), ("solve", "plain", False), ("pull",)]
    _check(_run_sequence(n, bc, dx, 2, steps, prec=prec), TOL[prec])


def test_repeat_solve_after_smooth3d_and_pulls():
    n, bc, dx = CASES3[0]
    b = random_rhs(n, 71)
    steps = [("push", b), ("smooth", 0.15), ("pull",), ("solve", "plain", False), ("smooth", 0.1),
             ("solve", "curl", True), ("pull",), ("solve", "plain", False)]
    _check(_run_sequence(n, bc, dx, 2, steps), 1e-12)


@pytest.mark.parametrize("bc", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_repeat_solve_2d(bc):
    n, dx = (12, 10), (0.1, 0.11)
    rng = np.random.default_rng(81)
    b = [rng.standard_normal((n[1], n[0])) for _ in range(3)]
    steps = [("push", b), ("solve", "plain", False), ("solve", "plain", False), ("pull",),
             ("solve", "grad", False), ("solve", "curl", False)]
    _check(_run_sequence(n, bc, dx, 4, steps, dim=2), 1e-12)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_async_solve_then_host_pull_and_push(prec):
    """ps_solve3d_async, then host pull / push at once: the transfers run after
    the solve on the context's stream."""
    n, bc, dx, reg = (64, 48, 40), (0, 0, 0), (0.05, 0.05, 0.05), 2
    announce("async ordering", prec)
    b1, b2 = random_rhs(n, 91, 1)[0], random_rhs(n, 92, 1)[0]
    ref = P.poisson_solver(PREC[prec])
    ref.setup3d(n, bc, dx, reg)
    want = []
    for b in (b1, b2):
        ref.push(b)
        ref.solve3d()
        w = np.empty_like(b)
        ref.pull(w)
        want.append(w)
    ref.close()
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, reg)
    s.push(b1)
    s.solve3d_async()
    got = np.full_like(b1, np.nan)
    s.pull(got)
    assert s.last_status == 0
    assert np.array_equal(got, want[0])
    s.solve3d_async()
    s.push(b2)                      # replaces the held rhs after the queued solve read it
    s.solve3d()
    s.pull(got)
    assert np.array_equal(got, want[1])
    s.close()


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_host_push_converts_like_device_push(prec):
    """The host-pointer push converts FP64 to FP32 on the device, rounding as
    a host conversion does: its solve equals a device push of the rounded
    data bit for bit, and pull widens exactly."""
    n, bc, dx = (32, 24, 512), (0, 0, 0), (0.1, 0.1, 1 / 256)
    b = random_rhs(n, 93, 1)[0]
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, 2)
    s.push(b)
    rb = np.empty_like(b)
    s.pull(None, None, None, rb)
    assert np.array_equal(rb, b.astype(np.float32).astype(np.float64))
    s.solve3d()
    host = np.empty_like(b)
    s.pull(host)
    d_in, d_out = DeviceBuffer(b.size * 4), DeviceBuffer(b.size * 4)
    d_in.upload(b.astype(np.float32))
    s.push_device(d_in.address)
    s.solve3d()
    s.pull_device(d_out.address)
    assert s.last_status == 0
    assert np.array_equal(host, d_out.download(np.float32, b.shape).astype(np.float64))
    s.close()
    d_in.free()
    d_out.free()


def test_device_memory_limit_gives_enomem_and_recovers():
    n, bc, dx = (64, 64, 64), (0, 0, 0), (1 / 32,) * 3
    announce("memory limit")
    b = random_rhs(n, 95, 1)[0]
    s = P.poisson_solver(P.PS_FP64)
    assert s.set_device_memory_limit(1 << 20)
    s.setup3d(n, bc, dx, 2)
    assert s.last_status == _capi.PS_ENOMEM
    assert s.device_memory_used() == 0
    s.push(b)
    assert s.last_status == _capi.PS_ESTATE
    # the set-up fits, the work buffers of the first solve do not
    s.set_device_memory_limit(16 << 20)
    s.setup3d(n, bc, dx, 2)
    assert s.last_status == 0
    used = s.device_memory_used()
    assert 0 < used <= 16 << 20
    s.push(b)
    s.solve3d()
    assert s.last_status == _capi.PS_ENOMEM
    s.set_device_memory_limit(0)
    s.solve3d()
    assert s.last_status == 0
    got = np.empty_like(b)
    s.pull(got)
    want, _ = O.solve(n, bc, dx, 2, [b])
    assert O.rel_l2(got, want[0]) <= 1e-12
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_non_finite_source_gives_non_finite_solution(prec, bad):
    n, bc, dx = (8, 8, 512), (0, 0, 0), (0.1, 0.1, 1 / 256)   # mixed: matrix-core z pass
    b = random_rhs(n, 97, 1)[0]
    b[100, 3, 4] = bad
    s = P.poisson_solver(PREC[prec])
    s.setup3d(n, bc, dx, 2)
    if prec == "mixed":
        assert s.zpass_variant() == "matrix-core"
    s.push(b)
    s.solve3d()
    assert s.last_status == 0
    got = np.zeros_like(b)
    s.pull(got)
    assert not np.isfinite(got).any()
    s.close()


def test_batch_member_set_up_again_and_pushed_directly():
    """A member set up again and pushed through its own context (not the
    batch's push): the batch notices before replaying and solves it alone."""
    announce("batch member re-setup, direct push")
    probs = [((12, 10, 9), (0, 0, 0), (0.1, 0.1, 0.1), 2)] * 3 + [((8, 8, 8), (0, 0, 0), (0.1, 0.1, 0.1), 2)]
    bt = P.batch(probs, P.PS_FP64)
    srcs = [random_rhs(p[0], 900 + i, 1)[0] for i, p in enumerate(probs)]
    for i, src in enumerate(srcs):
        bt.push(i, src)
    for _ in range(2):
        bt.solve()
    c = ctypes.c_void_p()
    assert bt._lib.ps_batch_context(bt._b, 1, ctypes.byref(c)) == 0
    n, bc, dx, reg = probs[1]
    assert bt._lib.ps_setup3d(c, (ctypes.c_int * 3)(*n), (ctypes.c_int * 3)(*bc), (ctypes.c_double * 3)(*dx),
                              reg) == 0
    new = random_rhs(n, 999, 1)[0]
    assert bt._lib.ps_push(c, new.ctypes.data_as(ctypes.c_void_p), None, None) == 0
    srcs[1] = new
    for _ in range(2):
        bt.solve()
    for i, (n, bc, dx, reg) in enumerate(probs):
        got = [np.full((n[2], n[1], n[0]), np.nan) for _ in range(3)]
        bt.pull(i, *got)
        ref, _ = O.solve(n, bc, dx, reg, [srcs[i]])
        assert O.rel_l2(got[0], ref[0]) <= 1e-12, i
    bt.close()
