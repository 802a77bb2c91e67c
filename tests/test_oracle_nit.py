"""Pins the oracle's non-linear inertialess transform of the realisation
(NIT_transform::transform_realisation, src/nonlinear/nit_transform.cc:127-164).

The reference has no tests of its own here.  The checks: the Gauss-Legendre
rule against numpy, Owen's T against its closed forms and adaptive
quadrature, the two target CDFs against their densities (finite
differences) and special cases, the bisection against a brute-force Python
restatement, and the transforms that must be identities.
"""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.stats import norm

import pyoracle as O


def test_gauss_legendre_matches_numpy():
    x, w = O.gauss_legendre_half(10)
    X, W = np.polynomial.legendre.leggauss(20)
    np.testing.assert_allclose(x, X[10:], rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, W[10:], rtol=0, atol=1e-14)


@pytest.mark.parametrize("a,tol", [(0.3, 1e-15), (1.0, 1e-13), (2.0, 1e-8)])
def test_owen_t_at_zero(a, tol):
    assert O.owen_t(0.0, a) == pytest.approx(math.atan(a) / (2 * math.pi), abs=tol)


@pytest.mark.parametrize("h", [-1.7, 0.0, 0.4, 1.3, 2.5])
def test_owen_t_alpha_one(h):
    p = norm.cdf(h)
    assert O.owen_t(h, 1.0) == pytest.approx(0.5 * p * (1 - p), abs=1e-14)


@pytest.mark.parametrize("h,a", [(0.5, 0.7), (1.3, 1.5), (-0.8, 0.4), (2.0, 1.0)])
def test_owen_t_quadrature(h, a):
    ref, _ = integrate.quad(lambda x: math.exp(-0.5 * h * h * (1 + x * x)) / (1 + x * x), 0, a,
                            epsabs=1e-15)
    assert O.owen_t(h, a) == pytest.approx(ref / (2 * math.pi), abs=1e-10)


def test_gram_charlier_normal_case():
    for x in np.linspace(-4, 4, 17):
        assert O.gram_charlier_cdf(x, 0.0, 3.0) == pytest.approx(norm.cdf(x), abs=1e-12)


@pytest.mark.parametrize("s,k", [(0.5, 3.7), (-0.3, 2.4), (3.25, 2.4)])
def test_gram_charlier_density(s, k):
    """dF/dx = phi(x) (1 + s He3 / 6 + (k - 3) He4 / 24)."""
    h = 1e-5
    for x in np.linspace(-3, 3, 13):
        d = (O.gram_charlier_cdf(x + h, s, k) - O.gram_charlier_cdf(x - h, s, k)) / (2 * h)
        he3, he4 = x ** 3 - 3 * x, x ** 4 - 6 * x ** 2 + 3
        expect = norm.pdf(x) * (1 + s * he3 / 6 + (k - 3) * he4 / 24)
        assert d == pytest.approx(expect, abs=1e-7)


def test_skew_normal_cdf():
    for x in np.linspace(-3, 3, 7):
        assert O.skew_normal_cdf(x, 0.4, 1.7, 0.0) == pytest.approx(norm.cdf(x, 0.4, 1.7),
                                                                       abs=1e-15)
    # density 2 phi(z) Phi(alpha z) / sigma
    mu, sig, al, h = 0.2, 1.3, 1.5, 1e-5
    for x in np.linspace(-2.5, 3, 12):
        d = (O.skew_normal_cdf(x + h, mu, sig, al) - O.skew_normal_cdf(x - h, mu, sig, al)) / (2 * h)
        z = (x - mu) / sig
        assert d == pytest.approx(2 * norm.pdf(z) * norm.cdf(al * z) / sig, abs=1e-7)


def bisection_ref(F, y, a, b, eps, max_iter):
    """src/linalg.hh:199-213, F evaluated at a and b again in every step."""
    i = 0
    while True:
        c = 0.5 * (a + b)
        fc = F(c) - y
        if (F(a) - y) * fc < 0:
            b = c
        if (F(b) - y) * fc < 0:
            a = c
        i += 1
        if not (i < max_iter and (b - a) > eps and abs(fc) > eps):
            return c


@pytest.mark.parametrize("dist", [O.GRAM_CHARLIER, O.SKEW_NORMAL])
def test_nit_bisection_restatement(dist):
    rng = np.random.default_rng(4)
    data = rng.standard_normal(40) * 1.3
    params = dict(skewness=0.4, kurtosis=3.5, mean=0.1, sigma=1.2, alpha=1.5)
    got = O.nit_realisation(data, 1.3, dist, interval=(-5.0, 5.0), eps=1e-7, max_iter=60,
                            **params)
    if dist == O.GRAM_CHARLIER:
        def F(x):
            return O.gram_charlier_cdf(x, params["skewness"], params["kurtosis"])
    else:
        def F(x):
            return O.skew_normal_cdf(x, params["mean"], params["sigma"], params["alpha"])
    for v, g in zip(data, got):
        y = 0.5 * math.erfc(-(v / 1.3) / math.sqrt(2))
        assert g == bisection_ref(F, y, -5.0, 5.0, 1e-7, 60)


def test_nit_identity_transforms():
    data = np.linspace(-2.5, 2.5, 21)
    sd = 1.7
    gc = O.nit_realisation(data, sd, O.GRAM_CHARLIER, skewness=0.0, kurtosis=3.0,
                           interval=(-6.0, 6.0), eps=1e-12, max_iter=200)
    np.testing.assert_allclose(gc, data / sd, atol=1e-10)
    sn = O.nit_realisation(data, sd, O.SKEW_NORMAL, mean=0.3, sigma=0.8, alpha=0.0,
                           interval=(-6.0, 6.0), eps=1e-12, max_iter=200)
    np.testing.assert_allclose(sn, 0.3 + 0.8 * data / sd, atol=1e-10)


# ---- the reference's own known answers (src/test/owen-t-test.cc,
# src/test/skew-normal-test.cc), at their tolerance 1e-5 ----

def test_owen_t_reference_cases():
    assert O.owen_t(1.0, 0.0) == pytest.approx(0.0, abs=1e-5)
    assert O.owen_t(1.0, 1.0) == pytest.approx(O.owen_t(-1.0, 1.0), abs=1e-5)
    assert O.owen_t(1.0, -1.0) == pytest.approx(-O.owen_t(1.0, 1.0), abs=1e-5)
    for h, a, v in [(2.0, 0.5, 0.00862508), (1.0, 1.5, 0.07573), (1.5, 2.0, 0.0333832),
                    (1.0, 0.1, 0.00960528)]:
        assert O.owen_t(h, a) == pytest.approx(v, abs=1e-5)


def test_skew_normal_reference_cases():
    expect = [-1.0842021724855044e-19, 5.596621256709344e-13, 3.143618040497842e-7,
              0.0017188799452888537, 0.14758361765043326, 0.6844083720823747,
              0.9545000504654456, 0.9973002039372995, 0.9999366575163338]
    for x, v in zip(range(-4, 5), expect):
        assert O.skew_normal_cdf(float(x), 0.0, 1.0, 2.0) == pytest.approx(v, abs=1e-5)
