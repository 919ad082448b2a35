"""The product's pipeline design (R2C along x, real octant of
Ghat*zeta/N, Hermitian-part operator coefficients, lazily computed filtered
rhs) reproduces the oracle's C2C restatement of the reference to round-off,
for every boundary tuple, operator, regularisation order and reprojection,
including odd periodic lengths and random (non band-limited) sources whose
Nyquist content is large."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.r2c_model import model_solve

CASES = [
    ((6, 5, 4), (0, 0, 0)),
    ((8, 7, 6), (1, 1, 0)),
    ((7, 6, 5), (1, 1, 0)),
    ((6, 8, 5), (0, 0, 1)),
    ((5, 6, 7), (1, 1, 1)),
    ((6, 6, 6), (1, 1, 1)),
]


@pytest.mark.parametrize("n,bc", CASES)
@pytest.mark.parametrize("op", ["plain", "grad", "div", "curl"])
@pytest.mark.parametrize("reg,reproject", [(0, False), (2, True), (5, False), (6, True)])
def test_model_matches_oracle(n, bc, op, reg, reproject):
    rng = np.random.default_rng(hash((n, bc, op, reg)) % 2 ** 32)
    dx = (0.13, 0.1, 0.17)
    b = [rng.standard_normal((n[2], n[1], n[0])) for _ in range(3)]
    lo, ro = O.solve(n, bc, dx, reg, b, op=op, reproject=reproject)
    lm, rm = model_solve(n, bc, dx, reg, b, op=op, reproject=reproject, want_rhs=True)
    for d in range(3):
        assert (lo[d] is None) == (lm[d] is None)
    got = [x for x in lm if x is not None]
    ref = [x for x in lo if x is not None]
    assert O.rel_l2(got, ref) < 1e-13
    assert O.rel_l2(rm, ro) < 1e-13
