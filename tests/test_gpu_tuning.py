"""Every tile shape / A2 layout the tuning knobs can select gives the same
FP64 result as the oracle (rel-L2 <= 1e-12)."""
import os

import pytest

from oracle import oracle as O
from tests.gpu_util import announce, gpu_solve, random_rhs

pytestmark = pytest.mark.gpu

VARIANTS = [
    {"PS_TUNE_ZBLK": "0"},
    {"PS_TUNE_ZBLK": "-1"},
    {"PS_TUNE_ZBLK": "2"},
    {"PS_TUNE_ZBLK": "3", "PS_TUNE_XCD": "0"},
    {"PS_TUNE_BY": "2", "PS_TUNE_BZ": "1", "PS_TUNE_PX": "1"},
    {"PS_TUNE_E": "32"},
    {"PS_TUNE_E": "8", "PS_TUNE_LDS_KB": "32"},
    {"PS_TUNE_STATIC": "0"},
    {"PS_TUNE_STATIC3": "1"},
    {"PS_TUNE_TWREG_Z": "1", "PS_TUNE_TWREG_Y": "1"},
    {"PS_TUNE_OCC4_Z": "1"},
    {"PS_TUNE_HALFB_Z": "1", "PS_TUNE_HALFB_Y": "1"},
    {"PS_TUNE_ZWAVE": "0"},
    {"PS_TUNE_ZWAVE": "1", "PS_TUNE_HALFB_Z": "0"},
]


@pytest.mark.parametrize("env", VARIANTS, ids=lambda e: ",".join("%s=%s" % kv for kv in e.items()))
@pytest.mark.parametrize("case", [((24, 20, 18), (0, 0, 0), "curl"), ((32, 16, 20), (1, 1, 0), "plain"),
                                  ((32, 64, 32), (0, 0, 0), "plain"), ((32, 32, 64), (0, 0, 0), "curl"),
                                  ((64, 128, 64), (1, 1, 0), "grad")],
                         ids=["uuu-curl", "ppu-plain", "uuu-static-plain", "uuu-static-curl", "ppu-static-grad"])
def test_variant_parity(env, case, monkeypatch):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    n, bc, op = case
    dx = (0.1, 0.11, 0.09)
    announce(env, n, bc, op)
    b = random_rhs(n, 23)
    lo, _ = O.solve(n, bc, dx, 2, b, op=op)
    lg, _ = gpu_solve(n, bc, dx, 2, b, op=op, prec="fp64")
    got = [lg[d] for d in range(3) if lo[d] is not None]
    ref = [lo[d] for d in range(3) if lo[d] is not None]
    assert O.rel_l2(got, ref) <= 1e-12


def test_planner_rules_resolve_as_documented():
    """The planner's size rules, read back from the tuning a setup used:
    half-width z tiles for z transforms up to 512, register y twiddles for
    single-precision y transforms from 1024 (never FP64), the fused z kernel
    for FP64 pencils of 1024 and the wavefront kernel for FP32 ones."""
    import poissonsolver_amd as P
    cases = [
        # (ncell, precision, halfb_z, twreg_y, z variant)
        ((6, 512, 8), P.PS_FP32, 1, 1, None),        # My 1024, Mz 16
        ((6, 512, 8), P.PS_FP64, 1, 0, None),
        ((6, 256, 8), P.PS_FP32, 1, 0, None),        # My 512
        ((6, 8, 512), P.PS_FP32, 0, 0, "pencil-per-wavefront"),   # Mz 1024
        ((6, 8, 512), P.PS_FP64, 0, 0, "specialised"),
        ((6, 8, 256), P.PS_FP64, 1, 0, "pencil-per-wavefront"),   # Mz 512
    ]
    for n, prec, hb, ty, zv in cases:
        s = P.poisson_solver(prec)
        s.setup3d(n, (0, 0, 0), (0.1, 0.1, 0.1), 2)
        t = s.tuning()
        assert (t["halfb_z"], t["twreg_y"]) == (hb, ty), (n, prec, t)
        if zv is not None:
            assert s.zpass_variant() == zv, (n, prec, s.zpass_variant())
        s.close()
