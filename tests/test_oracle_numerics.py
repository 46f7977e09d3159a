"""Pins of the oracle's numerics (CPU): special functions against mpmath
fixtures, QAGS against scipy's QUADPACK (the same published dqagse) and closed
forms, the spline and the reference's Simpson rule."""
import math
import os

import numpy as np
import pytest
import scipy.integrate as si

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def kat():
    return np.load(os.path.join(GOLD, "bessel_kat.npz"))


def test_k0_k1_vs_mpmath(kat):
    for x, k0, k1 in zip(kat["xk"], kat["k0"], kat["k1"]):
        assert abs(O.bessel_k0(x) / k0 - 1) < 2e-15, x
        assert abs(O.bessel_k1(x) / k1 - 1) < 2e-15, x


def test_j1_vs_mpmath(kat):
    for x, j in zip(kat["xj"], kat["j1"]):
        env = 1.0 if x < 2 else math.sqrt(2 / (math.pi * x))
        assert abs(O.bessel_j1(x) - j) <= 1e-15 * env, x


def test_as_k1_within_published_bound(kat):
    # A&S 9.8.7/9.8.8: |eps| < 8e-9 (x <= 2, on x K1) and 2.2e-7 (x > 2, on sqrt(x) e^x K1)
    for x, k1 in zip(kat["xk"], kat["k1"]):
        if x < 1e-3 or x > 300:
            continue
        v = O.bessel_k1_as(x)
        if x <= 2:
            assert abs(x * v - x * k1) < 8e-9 * 1.01, x
        else:
            assert abs(v / k1 - 1) * 1.25331414 < 2.2e-7 * 1.01, x
    # the polynomial, not the exact function: its error is visible at the 1e-8 level
    errs = [abs(O.bessel_k1_as(x) / k - 1) for x, k in zip(kat["xk"], kat["k1"]) if 2 < x < 50]
    assert max(errs) > 1e-8


def test_i1_as_small_argument():
    # A&S 9.8.3: I1(x)/x = 1/2 + ... exactly 1/2 at x -> 0
    assert abs(O.bessel_i1_as(1e-8) / 1e-8 - 0.5) < 1e-15


QAGS_CASES = [
    (lambda x: math.sqrt(x) * math.log(x) if x > 0 else 0.0, 0.0, 1.0, -4.0 / 9.0),
    (lambda x: math.log(x) / math.sqrt(x) if x > 0 else 0.0, 0.0, 1.0, -4.0),
    (lambda x: math.exp(-x) * math.cos(30 * x), 0.0, 10.0, (1 - math.exp(-10) * (math.cos(300) - 30 * math.sin(300))) / 901.0),
    (lambda x: 1.0 / (1.0 + x * x), 0.0, 10.0, math.atan(10.0)),
    (lambda x: x ** 2 / (x ** 2 + 1e-6) * math.sin(20 * x), 0.0, 10.0, None),
]


@pytest.mark.parametrize("case", range(len(QAGS_CASES)))
@pytest.mark.parametrize("eps", [1e-4, 1e-10])
def test_qags_matches_quadpack_trace(case, eps):
    f, a, b, exact = QAGS_CASES[case]
    res, err, nev, ier, (al, bl, rl, el) = O.qags(f, a, b, eps, eps)
    s = si.quad(f, a, b, epsabs=eps, epsrel=eps, limit=1000, full_output=1)
    info = s[2]
    assert ier == 0
    assert nev == info["neval"]
    assert res == s[0]
    n = info["last"]
    assert np.array_equal(np.sort(al), np.sort(info["alist"][:n]))
    assert np.array_equal(np.sort(rl), np.sort(info["rlist"][:n]))
    if exact is not None:
        assert abs(res - exact) <= max(err, 1e-14)


def test_simpson_weights_as_written():
    # n even: weights 1,4,2,...,4,2,1 (src/UpcCrossSection.cpp:139-150)
    n, h = 200, 20.0 / 199
    v = np.zeros(n)
    for i in range(n):
        v[:] = 0
        v[i] = 1.0
        w = round(O.simpson(v, h) * 3 / h, 9)
        expect = 1 if i in (0, n - 1) else (4 if i % 2 == 1 else 2)
        assert w == expect, i
    assert O.simpson(np.ones(n), h) == pytest.approx(596 * h / 3, rel=1e-15)


def test_spline_matches_natural_cubic_spline():
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.uniform(0.1, 1.0, 40))
    y = np.sin(x) + 0.1 * x * x
    xq = np.concatenate([x, rng.uniform(x[0], x[-1], 300)])
    ref = CubicSpline(x, y, bc_type="natural")(xq)
    got = O.spline_eval(x, y, xq)
    assert np.max(np.abs(got - ref)) < 1e-13
    # outside the knot range the reference's GSL spline raises GSL_EDOM
    assert np.isnan(O.spline_eval(x, y, [x[0] - 1e-9])[0])
