"""validate_process (src/arma.cc:15-70): roots of 1 - phi_1 x - ... - phi_n x^n
over the row-major coefficients, all required outside the unit circle.

The reference calls GSL's gsl_poly_complex_solve (companion matrix, QR);
GSL is absent here, so the roots are pinned against numpy.roots (LAPACK
eigenvalues of the companion matrix) and against polynomials with known
roots.  Host code: no GPU needed."""
import numpy as np
import pytest

from arma_amd import _abi
import pyoracle as O


def poly_of(phi):
    c = -np.ascontiguousarray(phi, np.float64).reshape(-1)
    c[0] = 1.0
    return c


@pytest.mark.parametrize("deg", [1, 2, 3, 7, 20, 64, 150, 342])
def test_roots_match_numpy(deg):
    rng = np.random.default_rng(deg)
    phi = rng.standard_normal(deg + 1) * 0.5
    phi[0] = 0.0
    try:
        nbad, roots = _abi.validate_process(phi.reshape(1, 1, -1))
    except _abi.ArmaHipError as e:
        nbad, roots = e.n_bad, e.roots
    ref = np.roots(poly_of(phi)[::-1])
    assert len(roots) == deg
    np.testing.assert_allclose(np.sort(np.abs(roots)), np.sort(np.abs(ref)), rtol=1e-9,
                               atol=1e-12)
    # every root is a root: |p(r)| small relative to the terms' size
    c = poly_of(phi)
    for r in roots:
        terms = c * r ** np.arange(len(c))
        assert abs(terms.sum()) <= 1e-9 * np.abs(terms).sum()
    assert nbad == int(np.sum(~(np.abs(ref) > 1.0)))


def test_known_roots():
    # (1 - x/2)(1 - x/3)(1 + x/4) = 1 - phi_1 x - phi_2 x^2 - phi_3 x^3: roots 2, 3, -4
    c = np.polynomial.polynomial.polyfromroots([2.0, 3.0, -4.0])
    c = c / c[0]
    phi = -c
    phi[0] = 0.0
    nbad, roots = _abi.validate_process(phi.reshape(1, 1, 4))
    assert nbad == 0
    np.testing.assert_allclose(np.sort(roots.real), [-4.0, 2.0, 3.0], rtol=1e-13)
    np.testing.assert_allclose(roots.imag, 0.0, atol=1e-13)
    # a root inside the unit circle is rejected, as the reference's !(|r| > 1) does
    c = np.polynomial.polynomial.polyfromroots([0.9, 5.0])
    c = c / c[0]
    phi = -c
    phi[0] = 0.0
    with pytest.raises(_abi.ArmaHipError) as e:
        _abi.validate_process(phi.reshape(1, 1, 3))
    assert e.value.n_bad == 1
    assert "not stationary" in str(e.value)


def test_leading_zero_is_an_error():
    phi = np.zeros((2, 2, 2))
    phi[0, 0, 1] = 0.3
    with pytest.raises(_abi.ArmaHipError) as e:
        _abi.validate_process(phi)  # last coefficient 0: GSL refuses, so does this
    assert "Can not find roots" in str(e.value)


def test_single_coefficient_is_valid():
    assert _abi.validate_process(np.zeros((1, 1, 1)))[0] == 0


@pytest.mark.parametrize("kind,beta,order", [(O.STANDING, (0.8, 0.8), 7),
                                             (O.PROPAGATING, (-0.8, 0.0), 7)])
def test_models_against_numpy(kind, beta, order):
    acf = O.acf(kind, (10, 10, 10), velocity=0.8, beta=beta)
    phi, _ = O.yw_choi(acf, order)
    ref = np.roots(poly_of(phi)[::-1])
    want = int(np.sum(~(np.abs(ref) > 1.0)))
    try:
        nbad, roots = _abi.validate_process(phi)
    except _abi.ArmaHipError as e:
        nbad, roots = e.n_bad, e.roots
    np.testing.assert_allclose(np.sort(np.abs(roots)), np.sort(np.abs(ref)), rtol=1e-7)
    assert nbad == want
