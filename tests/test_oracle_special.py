"""Pins the oracle's special functions (oracle/special.c).

* Golden vectors: outputs of the reference's own special_functions.cpp
  (tests/golden/special_ref.npz, written by tests/golden/make_special_golden.py
  from oracle/_ref).  Same algorithm, same constants, same libm -> bitwise.
* Independent mathematics: scipy.special.sici (Si) and mpmath quadrature of
  (1 - J0(t))/t (Ji0), with the accuracy the published methods promise.
"""
import os

import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "special_ref.npz")


def test_sine_int_matches_reference_golden_bitwise():
    g = np.load(GOLD)
    y = O.sine_int(g["si_x"])
    assert np.array_equal(y, g["si_y"])


def test_bessel_int_J0_matches_reference_golden_bitwise():
    g = np.load(GOLD)
    y = O.bessel_int_J0(g["ji0_x"])
    assert np.array_equal(y, g["ji0_y"])


def test_sine_int_against_scipy():
    x = np.concatenate([np.linspace(1e-6, 30, 3000), np.geomspace(30, 1e4, 500)])
    si, _ = sp.sici(x)
    rel = np.abs(O.sine_int(x) - si) / np.abs(si)
    # NR cisi with eps = 1e-10: worst near the branch point t = 2
    assert rel.max() < 3e-11
    assert rel[x > 100].max() < 1e-14
    # odd symmetry and Si(0) = 0
    assert np.array_equal(O.sine_int(-x), -O.sine_int(x))
    assert O.sine_int(np.zeros(1))[0] == 0.0


def test_bessel_int_J0_against_quadrature():
    mpmath = pytest.importorskip("mpmath")
    xs = [1e-3, 0.5, 1.0, 3.0, 7.9, 8.0, 8.1, 15.0, 40.0, 200.0]
    for x in xs:
        exact = float(mpmath.quad(lambda t: (1 - mpmath.besselj(0, t)) / t,
                                  mpmath.linspace(0, x, 40)))
        got = O.bessel_int_J0(np.array([x]))[0]
        assert abs(got - exact) <= 1e-14 + 2e-13 * abs(exact), (x, got, exact)
    # Ji0(0) = 0 and the large-x form ln(x/2) + gamma
    assert abs(O.bessel_int_J0(np.array([0.0]))[0]) < 1e-15
    big = O.bessel_int_J0(np.array([1e4]))[0]
    assert abs(big - (np.log(5e3) + np.euler_gamma)) < 1e-5


def test_golden_covers_both_branches():
    g = np.load(GOLD)
    assert (g["si_x"] > 2).sum() > 100 and (np.abs(g["si_x"]) <= 2).sum() > 100
    assert (g["ji0_x"] >= 8).sum() > 100 and (g["ji0_x"] < 8).sum() > 100
