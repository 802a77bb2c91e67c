"""Pins the oracle's Yule-Walker solvers and AR recurrence.

Known answers are the reference's own test cases
(src/test/yule-walker-test.cc:141-234); the AR recurrence
(src/generator/ar_model_surf.cc:3-36) is pinned by identities that a dropped
term, a wrong sign or a transposed index would break.
"""
import itertools

import numpy as np
import pytest

import pyoracle as O


def exponential_acf(shape):
    a, b, c = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return 0.9 ** a * 0.88 ** b * 0.95 ** c


def standing_wave_test_acf(shape):
    # src/test/yule-walker-test.cc:25-47
    t, x, y = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    alpha, beta = 0.06, 0.8
    velocity = 2 * beta
    acf = np.exp(-alpha * (2 * t + x + y)) * np.cos(velocity * t) * np.cos(beta * x) * np.cos(0 * beta * y)
    return acf / acf[0, 0, 0]


@pytest.mark.parametrize("order,var,tol,gen", [
    (2, 239.2780, 1e-4, exponential_acf),
    (3, 239.2780, 1e-4, exponential_acf),
    (4, 239.2780, 1e-4, exponential_acf),
    (7, 2.0, 1e-2, standing_wave_test_acf),
])
def test_choi_equals_gauss(order, var, tol, gen):
    acf = var * gen((order + 1,) * 3)
    choi, _ = O.yw_choi(acf, order, determine_order=False, chop=False)
    gauss = O.yw_gauss(acf, (order,) * 3)
    assert choi.shape == gauss.shape == (order,) * 3
    assert np.max(np.abs(choi - gauss)) < tol


def yw_residual(acf, phi):
    """max |gamma(l) - sum_k phi(k) gamma(l-k)| over lags l != 0 of phi's box."""
    p = phi.shape
    worst = 0.0
    for l in itertools.product(*[range(n) for n in p]):
        if l == (0, 0, 0):
            continue
        s = 0.0
        for k in itertools.product(*[range(n) for n in p]):
            d = tuple(abs(l[i] - k[i]) for i in range(3))
            s += phi[k] * acf[d]
        worst = max(worst, abs(acf[l] - s))
    return worst


def test_gauss_solves_normal_equations():
    acf = 239.278 * exponential_acf((4, 4, 4))
    phi = O.yw_gauss(acf, (3, 3, 3))
    assert phi[0, 0, 0] == 0.0
    assert yw_residual(acf, phi) < 1e-9


@pytest.mark.parametrize("var_eps,shape", [(1e-2, (3, 3, 3)), (1e-3, (4, 4, 4)), (1e-4, (5, 5, 5))])
def test_choi_converged_solves_normal_equations(var_eps, shape):
    """When the variance converges, Choi returns the exact order-l solution."""
    acf = 2.0 * standing_wave_test_acf((8, 8, 8))
    phi, _ = O.yw_choi(acf, 7, var_eps=var_eps, determine_order=True, chop=False)
    assert phi.shape == shape
    assert yw_residual(acf, phi) < 1e-12


@pytest.mark.parametrize("M", [3, 4, 5])
def test_choi_at_max_order_truncates_next_stage(M):
    """At max order the reference maps the newest stage (lags up to M)
    into an M^3 array (src/yule_walker.cc:521-527): the result equals the
    (M+1)^3 Gauss solution cut to M^3."""
    acf = 2.0 * standing_wave_test_acf((M + 2,) * 3)
    phi, _ = O.yw_choi(acf, M, determine_order=False, chop=False)
    g = O.yw_gauss(acf, (M + 1,) * 3)
    assert np.max(np.abs(phi - g[:M, :M, :M])) < 1e-11


def test_exponential_acf_is_separable_ar1():
    # gamma = 0.9^a 0.88^b 0.95^c is a separable AR(1): phi has (2,2,2)
    # support with phi(1,1,1) = -0.9*0.88*0.95 etc.
    acf = exponential_acf((4, 4, 4))
    phi = O.yw_gauss(acf, (2, 2, 2))
    r = np.array([0.9, 0.88, 0.95])
    expect = np.zeros((2, 2, 2))
    for k in itertools.product((0, 1), repeat=3):
        if k != (0, 0, 0):
            expect[k] = -np.prod([(-r[d]) ** k[d] for d in range(3)])
    np.testing.assert_allclose(phi, expect, atol=1e-12)


def test_determine_the_order():
    acf = 239.278 * exponential_acf((11, 11, 11))
    phi, _ = O.yw_choi(acf, 10, var_eps=1e-5, determine_order=True, chop=False)
    assert phi.shape == (2, 2, 2)


def test_chop():
    acf = 239.278 * exponential_acf((11, 11, 11))
    phi, _ = O.yw_choi(acf, 10, var_eps=1e-5, determine_order=False, chop=True)
    assert phi.shape == (2, 2, 2)


def test_real_case():
    acf = 2.0 * standing_wave_test_acf((11, 11, 11))
    phi, _ = O.yw_choi(acf, 10, var_eps=1e-5)
    assert phi.shape[2] == 2


def test_chop_right_semantics():
    phi = np.zeros((4, 4, 4))
    phi[1, 2, 1] = 0.5
    phi[0, 0, 1] = 0.25
    out = (O.C.c_int * 3)()
    O.lib().orc_chop_right(phi.reshape(-1), O._i3(phi.shape), 1e-10, out)
    assert tuple(out) == (2, 3, 2)


def test_closed_form_acf():
    shp = (4, 5, 6)
    delta = (1.0, 0.5, 2.0)
    got = O.acf(O.PROPAGATING, shp, delta=delta, amplitude=1.5, velocity=0.8,
                alpha=(0.06, 0.1, 0.2), beta=(-0.8, 0.3))
    i, j, k = np.meshgrid(*[np.arange(n) for n in shp], indexing="ij")
    ref = 1.5 * np.exp(-(0.06 * i * 1.0 + 0.1 * j * 0.5 + 0.2 * k * 2.0)) * \
        np.cos(0.8 * i * 1.0 - 0.8 * j * 0.5 + 0.3 * k * 2.0)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)
    got = O.acf(O.STANDING, shp, delta=delta, velocity=0.8, beta=(0.8, 0.8))
    ref = np.exp(-0.06 * (i * 1.0 + j * 0.5 + k * 2.0)) * np.cos(0.8 * i) * np.cos(0.8 * j * 0.5) * \
        np.cos(0.8 * k * 2.0)
    np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_white_noise_variance():
    acf = exponential_acf((3, 3, 3))
    phi = np.arange(8, dtype=float).reshape(2, 2, 2) / 10
    assert O.white_noise_variance(acf, phi) == pytest.approx(acf[0, 0, 0] - np.sum(phi * acf[:2, :2, :2]),
                                                              abs=1e-15)


def test_partition_shape():
    assert O.partition_shape((16, 32, 32), (1024, 256, 256), (7, 7, 7), 8) == (16, 32, 32)
    # guess: g1 = max(2*order,10), g2 = ceil(shape/(7*cbrt(par))), min + |g1-g2|/2
    assert O.partition_shape((0, 0, 0), (100, 100, 100), (7, 7, 7), 8) == (11, 11, 11)


# ---- AR recurrence ----

def brute_ar(eps, phi):
    z = eps.copy()
    p = phi.shape
    for t, x, y in itertools.product(*[range(n) for n in z.shape]):
        s = 0.0
        for k in range(min(t + 1, p[0])):
            for i in range(min(x + 1, p[1])):
                for j in range(min(y + 1, p[2])):
                    s += phi[k, i, j] * z[t - k, x - i, y - j]
        z[t, x, y] += s
    return z


def test_ar_brute_force():
    rng = np.random.default_rng(1)
    eps = rng.standard_normal((5, 6, 7))
    phi = rng.standard_normal((2, 3, 2)) * 0.1
    phi[0, 0, 0] = 0.0
    np.testing.assert_array_equal(O.ar_generate(eps, phi), brute_ar(eps, phi))


def test_ar_inverse_filter_recovers_noise():
    rng = np.random.default_rng(2)
    eps = rng.standard_normal((12, 10, 9))
    phi = rng.standard_normal((3, 3, 3)) * 0.05
    phi[0, 0, 0] = 0.0
    z = O.ar_generate(eps, phi)
    # eps(t,x,y) = zeta - sum_{k,i,j} phi * zeta(t-k, x-i, y-j), zero outside
    zp = np.zeros((z.shape[0] + 2, z.shape[1] + 2, z.shape[2] + 2))
    zp[2:, 2:, 2:] = z
    rec = z.copy()
    for k, i, j in itertools.product(range(3), repeat=3):
        rec -= phi[k, i, j] * zp[2 - k:2 - k + z.shape[0], 2 - i:2 - i + z.shape[1], 2 - j:2 - j + z.shape[2]]
    np.testing.assert_allclose(rec, eps, atol=1e-12)


def test_ar_impulse_response():
    eps = np.zeros((10, 3, 3))
    eps[0, 1, 1] = 1.0
    phi = np.zeros((2, 1, 1))
    phi[1, 0, 0] = 0.7
    z = O.ar_generate(eps, phi)
    np.testing.assert_allclose(z[:, 1, 1], 0.7 ** np.arange(10), rtol=1e-15)
    assert np.count_nonzero(z) == 10
    # an in-row tap propagates along y only
    phi = np.zeros((1, 1, 2))
    phi[0, 0, 1] = -0.5
    eps = np.zeros((1, 1, 8))
    eps[0, 0, 0] = 1.0
    z = O.ar_generate(eps, phi)
    np.testing.assert_allclose(z[0, 0], (-0.5) ** np.arange(8), rtol=1e-15)


def test_ar_chunked_equals_one_shot():
    rng = np.random.default_rng(3)
    eps = rng.standard_normal((16, 8, 8))
    phi = rng.standard_normal((3, 2, 3)) * 0.05
    phi[0, 0, 0] = 0
    one = O.ar_generate(eps, phi)
    z = eps.copy()
    for t0 in range(0, 16, 5):
        t1 = min(t0 + 4, 15)
        z = O.ar_generate(z, phi, lo=(t0, 0, 0), hi=(t1, 7, 7))
    np.testing.assert_array_equal(z, one)
    dag = O.ar_generate_dag(eps, phi, (4, 3, 3))
    np.testing.assert_array_equal(dag, one)


def test_gsl_variance():
    assert O.gsl_variance(np.array([1.0, 2.0, 3.0, 4.0])) == pytest.approx(5.0 / 3.0, rel=1e-15)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(10001) * 3 + 7
    ref = float(np.var(x.astype(np.longdouble), ddof=1))
    assert O.gsl_variance(x) == pytest.approx(ref, rel=1e-14)
    xf = x.astype(np.float32)
    assert O.gsl_variance(xf) == pytest.approx(float(np.var(xf.astype(np.float64), ddof=1)), rel=1e-6)


def test_rescale_reads_contiguous_run():
    rng = np.random.default_rng(5)
    z = rng.standard_normal((5, 6, 7))
    start = (5 // 2) * 42 + (6 // 2) * 7 + 7 // 2
    cnt = (5 - 2) * (6 - 3) * (7 - 3)
    run = z.reshape(-1)[start:start + cnt]
    assert O.rescale_factor(z, 2.0) == pytest.approx(np.sqrt(2.0 / np.var(run, ddof=1)), rel=1e-14)
    # and it is NOT the sub-block variance
    sub = z[2:, 3:, 3:].reshape(-1)
    assert abs(np.var(sub, ddof=1) - np.var(run, ddof=1)) > 1e-6


def test_openmp_variants_equal_sequential():
    """The all-core CPU baseline's OpenMP entry points give the sequential
    oracle's results bit for bit (partitions and levels are independent)."""
    from arma_amd import _abi
    params = [O.MtParams(*p.as_tuple()) for p in _abi.dcmt_p521_params(64)]
    shape, part = (20, 24, 40), (8, 8, 16)
    n = 3 * 3 * 3
    seeds = np.arange(n, dtype=np.uint32) + 5
    a = O.noise_partitioned(params[:n], seeds, shape, part, 0.7)
    b = O.noise_partitioned(params[:n], seeds, shape, part, 0.7, threads=4)
    np.testing.assert_array_equal(a, b)
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, 5)
    np.testing.assert_array_equal(O.ar_generate_dag(a, phi, part),
                                  O.ar_generate_dag(a, phi, part, threads=4))
    np.testing.assert_array_equal(O.ar_generate_dag(a, phi, part), O.ar_generate(a, phi))
    z = O.ar_generate(a, phi)[:, :16, :32]
    args = ((2 * np.pi,) * 3, 12.0, (0.0, -10.0), (5.0, 0.0), (6, 5))
    np.testing.assert_array_equal(O.linear_potential(z, *args),
                                  O.linear_potential_omp(z, *args, 4))
