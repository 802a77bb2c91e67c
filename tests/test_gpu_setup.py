"""The FP64 model setup on the GPU: many Choi recursions in one launch.

Bar: every model of the batched launch gives the single solve's result (and
the host solver's) to the last bit: order, coefficients and the solver's
variance; a model whose recursion fails reports its own status without
affecting the others.
"""
import numpy as np
import pytest

from arma_amd import _abi
import pyoracle as O

pytestmark = pytest.mark.gpu


def models():
    out = []
    rng = np.random.default_rng(5)
    for order in (3, 5, 7, 9, 4, 7):
        a = float(rng.uniform(0.04, 0.1))
        b = float(rng.uniform(0.5, 1.2))
        kind = O.STANDING if order % 2 else O.PROPAGATING
        beta = (b, 0.7) if kind == O.STANDING else (-b, 0.0)
        out.append((O.acf(kind, (10, 10, 10), velocity=b, beta=beta, alpha=(a, a, a)), order))
    # a smaller ACF cube and the benchmark's standing wave
    out.append((O.acf(O.STANDING, (6, 6, 6), velocity=0.8, beta=(0.8, 0.8)), 5))
    out.append((O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8)), 7))
    return out


def test_choi_many_equals_single_and_host(hip_ctx):
    ms = models()
    many = hip_ctx.yule_walker_choi_many([a for a, _ in ms], [o for _, o in ms])
    for (acf, order), got in zip(ms, many):
        assert got is not None
        single = hip_ctx.yule_walker_choi(acf, order)
        host = _abi.yule_walker_choi(acf, order)
        for ref in (single, host):
            assert got[0].shape == ref[0].shape
            np.testing.assert_array_equal(got[0], ref[0])
            assert got[1] == ref[1]


def test_choi_many_failure_is_per_model(hip_ctx):
    """A zero-variance-free but singular ACF (all ones) makes R11 singular:
    that model fails, the others are unaffected."""
    good = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    bad = np.ones((10, 10, 10))
    res = hip_ctx.yule_walker_choi_many([good, bad, good], [7, 7, 5])
    assert res[1] is None
    np.testing.assert_array_equal(res[0][0], hip_ctx.yule_walker_choi(good, 7)[0])
    np.testing.assert_array_equal(res[2][0], hip_ctx.yule_walker_choi(good, 5)[0])
