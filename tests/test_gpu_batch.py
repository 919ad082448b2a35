"""Batched solves of uneven size (one hipGraph launch) and the tuning API.

Every problem of a batch must give the result of the same problem solved
alone, to the bit when both use the same tuning (the graph replays the same
kernels on the same buffers), and match the oracle: FP64 rel-L2 <= 1e-12,
FP32 <= 2e-6."""
import numpy as np
import pytest

import poissonsolver_amd as P
from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, random_rhs

pytestmark = pytest.mark.gpu

PROBLEMS = [
    # (ncell, bc, dx, reg)
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 2),
    ((24, 24, 24), (0, 0, 0), (1 / 12, 1 / 12, 1 / 12), 2),
    ((12, 10, 9), (0, 0, 0), (0.1, 0.12, 0.09), 0),
    ((16, 12, 20), (1, 1, 0), (1 / 16, 1 / 12, 0.1), 6),
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 4),
    ((14, 10, 7), (0, 0, 1), (0.11, 0.09, 1 / 7), 2),
]


def _pull(b, i, n, ncomp):
    shape = (n[2], n[1], n[0])
    out = [np.full(shape, np.nan) for _ in range(3)]
    b.pull(i, *out)
    return out[:ncomp]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6)])
def test_batch_matches_single_and_oracle(prec, tol):
    announce("batch plain", prec, len(PROBLEMS))
    b = P.batch(PROBLEMS, P.PS_FP64 if prec == "fp64" else P.PS_FP32)
    srcs = [random_rhs(p[0], 100 + i, 1)[0] for i, p in enumerate(PROBLEMS)]
    for i, s in enumerate(srcs):
        b.push(i, s)
    b.solve()                     # eager pass
    eager = [_pull(b, i, p[0], 1)[0] for i, p in enumerate(PROBLEMS)]
    b.solve()                     # capture + one graph launch
    graph = [_pull(b, i, p[0], 1)[0] for i, p in enumerate(PROBLEMS)]
    for i, (n, bc, dx, reg) in enumerate(PROBLEMS):
        assert np.array_equal(eager[i], graph[i]), i
        single = gpu_solve(n, bc, dx, reg, [srcs[i]], prec=prec)[0][0]
        assert np.array_equal(single, graph[i]), i
        ref = O.solve(n, bc, dx, reg, [srcs[i]])[0][0]
        assert O.rel_l2(graph[i], ref) <= tol, (i, O.rel_l2(graph[i], ref))
    # new data, same components: the captured graph is replayed on it
    srcs2 = [random_rhs(p[0], 200 + i, 1)[0] for i, p in enumerate(PROBLEMS)]
    for i, s in enumerate(srcs2):
        b.push(i, s)
    for rep in range(3):
        b.solve()
        for i, (n, bc, dx, reg) in enumerate(PROBLEMS):
            got = _pull(b, i, n, 1)[0]
            ref = O.solve(n, bc, dx, reg, [srcs2[i]])[0][0]
            assert O.rel_l2(got, ref) <= tol, (rep, i)
    b.close()


def test_batch_curl_fp64_and_rhs_pull():
    probs = [p for p in PROBLEMS if p[1] == (0, 0, 0)][:3]
    announce("batch curl fp64", len(probs))
    b = P.batch(probs, P.PS_FP64)
    b.set_ops(curl=True, reproject=True)
    srcs = [random_rhs(p[0], 300 + i, 3) for i, p in enumerate(probs)]
    for i, s in enumerate(srcs):
        b.push(i, *s)
    for _ in range(2):
        b.solve()
    for i, (n, bc, dx, reg) in enumerate(probs):
        shape = (n[2], n[1], n[0])
        lhs = [np.full(shape, np.nan) for _ in range(3)]
        rhs = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *lhs, *rhs)
        lo, ro = O.solve(n, bc, dx, reg, srcs[i], op="curl", reproject=True)
        assert O.rel_l2(lhs, lo) <= 1e-12
        assert O.rel_l2(rhs, ro) <= 1e-12
    b.close()


def test_batch_profile_and_autotune():
    probs = [((32, 32, 32), (0, 0, 0), (1 / 16,) * 3, 2), ((48, 48, 48), (0, 0, 0), (1 / 24,) * 3, 2),
             ((32, 32, 32), (0, 0, 0), (1 / 16,) * 3, 2)]
    announce("batch autotune fp32", len(probs))
    b = P.batch(probs, P.PS_FP32)
    b.autotune(iters=2)
    assert b.tuning(0) == b.tuning(2)
    srcs = [random_rhs(p[0], 400 + i, 1)[0] for i, p in enumerate(probs)]
    for i, s in enumerate(srcs):
        b.push(i, s)
    ms = b.profile(5)
    assert ms > 0
    b.solve()
    for i, (n, bc, dx, reg) in enumerate(probs):
        got = _pull(b, i, n, 1)[0]
        ref = O.solve(n, bc, dx, reg, [srcs[i]])[0][0]
        assert O.rel_l2(got, ref) <= 2e-6, i
    b.close()


def test_autotune_single_context_keeps_parity():
    n, bc, dx, reg = (32, 24, 20), (0, 0, 0), (0.06, 0.07, 0.08), 2
    announce("autotune single", n)
    s = P.poisson_solver(P.PS_FP32)
    s.setup3d(n, bc, dx, reg)
    res = s.autotune(iters=2)
    assert res is not None and res[1] > 0
    tuned = s.tuning()
    assert tuned == res[0]
    src = random_rhs(n, 500, 1)[0]
    s.push(src)
    s.solve3d()
    out = np.full((n[2], n[1], n[0]), np.nan)
    s.pull(out)
    assert s.last_status == 0
    ref = O.solve(n, bc, dx, reg, [src])[0][0]
    assert O.rel_l2(out, ref) <= 2e-6
    s.close()


@pytest.mark.parametrize("fields", [dict(zblk=-1), dict(zblk=0, zwave=0), dict(specialised=0, radix_e=8),
                                    dict(specialised=0, tile_y=4, tile_z=32, tile_x=2), dict(halfb_z=2, halfb_y=1)])
def test_set_tuning_parity_fp64(fields):
    n, bc, dx, reg = (32, 32, 32), (0, 0, 0), (1 / 16,) * 3, 2
    announce("set_tuning", fields)
    s = P.poisson_solver(P.PS_FP64)
    assert s.set_tuning(**fields)
    s.setup3d(n, bc, dx, reg)
    assert s.last_status == 0
    for k, v in fields.items():
        assert s.tuning()[k] == v
    src = random_rhs(n, 600, 1)[0]
    s.push(src)
    s.solve3d()
    out = np.full((n[2], n[1], n[0]), np.nan)
    s.pull(out)
    ref = O.solve(n, bc, dx, reg, [src])[0][0]
    assert O.rel_l2(out, ref) <= 1e-12
    s.close()


GROUPED = [
    # three of one grid, two of another, one alone (16^3 with another reg)
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 2),
    ((12, 10, 9), (0, 0, 0), (0.1, 0.12, 0.09), 0),
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 2),
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 4),
    ((12, 10, 9), (0, 0, 0), (0.1, 0.12, 0.09), 0),
    ((16, 16, 16), (0, 0, 0), (0.125, 0.125, 0.125), 2),
]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("mixed", 2e-6)])
@pytest.mark.parametrize("comps", ["x", "z", "xyz"])
def test_batch_groups_equal_single_solves(prec, tol, comps):
    """Problems of one grid solved as one multi-component solve give, to the
    bit, what each gives solved alone (component strides 3 n3 for one pushed
    component, n3 for three), and match the oracle."""
    code = {"fp64": P.PS_FP64, "fp32": P.PS_FP32, "mixed": P.PS_MIXED}[prec]
    announce("batch groups", prec, comps)
    b = P.batch(GROUPED, code)
    srcs = []
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        r = random_rhs(n, 500 + i, 3)
        rhs = {"x": [r[0], None, None], "z": [None, None, r[2]], "xyz": r}[comps]
        srcs.append(rhs)
        b.push(i, *rhs)
    for _ in range(3):
        b.solve()
        assert b.grouped() == 5
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        shape = (n[2], n[1], n[0])
        got = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *got)
        single, _ = gpu_solve(n, bc, dx, reg, srcs[i], prec=prec)
        ref, _ = O.solve(n, bc, dx, reg, srcs[i])
        for d in range(3):
            if srcs[i][d] is None:
                assert np.isnan(got[d]).all(), (i, d)
            else:
                assert np.array_equal(got[d], single[d]), (i, d)
        g = [got[d] for d in range(3) if srcs[i][d] is not None]
        rr = [ref[d] for d in range(3) if srcs[i][d] is not None]
        assert O.rel_l2(g, rr) <= tol, i
    b.close()


def test_batch_groups_fall_back():
    """A group whose members pushed different components, or whose operator
    is not plain, is solved member by member, with the same results."""
    announce("batch group fallback")
    b = P.batch(GROUPED, P.PS_FP64)
    srcs = [random_rhs(p[0], 600 + i, 3) for i, p in enumerate(GROUPED)]
    # problem 2 pushes X and Z, its group mates X only: that group falls back
    for i, r in enumerate(srcs):
        b.push(i, r[0], None, r[2] if i == 2 else None)
    for _ in range(2):
        b.solve()
    assert b.grouped() == 2                # the (12, 10, 9) pair still groups
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        rhs = [srcs[i][0], None, srcs[i][2] if i == 2 else None]
        shape = (n[2], n[1], n[0])
        got = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *got)
        ref, _ = O.solve(n, bc, dx, reg, rhs)
        g = [got[d] for d in range(3) if rhs[d] is not None]
        rr = [ref[d] for d in range(3) if rhs[d] is not None]
        assert O.rel_l2(g, rr) <= 1e-12, i
    # grad: not a plain operator, every problem on its own
    b.set_ops(grad=True)
    for i, r in enumerate(srcs):
        b.push(i, *r)
    for _ in range(2):
        b.solve()
    assert b.grouped() == 0
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        shape = (n[2], n[1], n[0])
        got = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *got)
        ref, _ = O.solve(n, bc, dx, reg, srcs[i], op="grad")
        assert O.rel_l2(got, ref) <= 1e-12, i
    b.close()


def test_batch_group_reproject_and_filtered_rhs():
    """A grouped plain solve with reprojection: each member's solution and the
    filtered, reprojected rhs it pulls match the oracle."""
    announce("batch group reproject")
    b = P.batch(GROUPED, P.PS_FP64)
    b.set_ops(reproject=True)
    srcs = [random_rhs(p[0], 700 + i, 3) for i, p in enumerate(GROUPED)]
    for i, r in enumerate(srcs):
        b.push(i, *r)
    for _ in range(2):
        b.solve()
    assert b.grouped() == 5
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        shape = (n[2], n[1], n[0])
        lhs = [np.full(shape, np.nan) for _ in range(3)]
        rhs = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *lhs, *rhs)
        lo, ro = O.solve(n, bc, dx, reg, srcs[i], reproject=True)
        assert O.rel_l2(lhs, lo) <= 1e-12, i
        assert O.rel_l2(rhs, ro) <= 1e-12, i
    b.close()


def test_batch_member_set_up_again_leaves_its_group():
    """A member set up again through its context returns to buffers of its
    own; its group is then solved member by member, with the same results."""
    import ctypes
    announce("batch member re-setup")
    b = P.batch(GROUPED, P.PS_FP32)
    c = ctypes.c_void_p()
    assert b._lib.ps_batch_context(b._b, 2, ctypes.byref(c)) == 0
    n, bc, dx, reg = GROUPED[2]
    st = b._lib.ps_setup3d(c, (ctypes.c_int * 3)(*n), (ctypes.c_int * 3)(*bc), (ctypes.c_double * 3)(*dx), reg)
    assert st == 0
    srcs = [random_rhs(p[0], 800 + i, 1)[0] for i, p in enumerate(GROUPED)]
    for i, s in enumerate(srcs):
        b.push(i, s)
    for _ in range(2):
        b.solve()
    assert b.grouped() == 2                # only the (12, 10, 9) pair
    for i, (n, bc, dx, reg) in enumerate(GROUPED):
        shape = (n[2], n[1], n[0])
        got = [np.full(shape, np.nan) for _ in range(3)]
        b.pull(i, *got)
        ref, _ = O.solve(n, bc, dx, reg, [srcs[i]])
        assert O.rel_l2(got[0], ref[0]) <= 2e-6, i
    b.close()
