"""GPU parity of the non-linear inertialess transform of the realisation
(arma_hip_nit_realisation) against the oracle.

Bar: the bisection makes the same sign decisions as the oracle except where
the target CDF minus Phi is within an ulp of zero or of the stopping bound,
so at least 99 % of the points are bit-identical in FP64 and every point is
within 1e-4 (a few bisection steps at the stopping width); FP32 results are
the FP64 roots of the FP32 inputs rounded to float, within 1e-4.
"""
import numpy as np
import pytest
import torch

import arma_amd
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

CASES = {
    # the reference's input/nit-standing-gramcharlier parameters
    "gram_charlier": dict(dist=O.GRAM_CHARLIER, skewness=3.25, kurtosis=2.4, interval=(-5.0, 5.0)),
    "gram_charlier_mild": dict(dist=O.GRAM_CHARLIER, skewness=0.4, kurtosis=3.6,
                               interval=(-5.0, 5.0)),
    # input/nit-standing-skewnormal
    "skew_normal": dict(dist=O.SKEW_NORMAL, mean=0.0, sigma=1.0, alpha=1.0, interval=(-5.0, 5.0)),
    "skew_normal_shifted": dict(dist=O.SKEW_NORMAL, mean=0.3, sigma=1.4, alpha=-2.0,
                                interval=(-6.0, 6.0)),
}


def run(ctx, data, gamma0, case, prec, eps=1e-6, max_iter=100):
    dt = torch.float64 if prec == arma_amd.FP64 else torch.float32
    z = torch.from_numpy(np.ascontiguousarray(data)).to(DEV, dt)
    kw = {k: v for k, v in case.items() if k != "dist"}
    if "sigma" in kw:
        kw["stdev"] = kw.pop("sigma")
    d = arma_amd.nit_desc(case["dist"], absolute_error=eps, max_iterations=max_iter, **kw)
    ctx.nit_realisation(z, data.shape, prec, gamma0, d)
    return z.cpu().numpy().astype(np.float64)


def oracle(data, gamma0, case, eps=1e-6, max_iter=100):
    kw = {k: v for k, v in case.items() if k != "dist"}
    return O.nit_realisation(data, np.sqrt(gamma0), case["dist"], eps=eps, max_iter=max_iter, **kw)


@pytest.mark.parametrize("name", sorted(CASES))
def test_nit_fp64_matches_oracle(hip_ctx, name):
    print(f"nit fp64 {name}", flush=True)
    rng = np.random.default_rng(11)
    gamma0 = 1.7
    data = rng.standard_normal((20, 33, 29)) * np.sqrt(gamma0)
    got = run(hip_ctx, data, gamma0, CASES[name], arma_amd.FP64)
    ref = oracle(data, gamma0, CASES[name])
    assert np.mean(got == ref) >= 0.99
    assert np.max(np.abs(got - ref)) <= 1e-4


def test_nit_fp32(hip_ctx):
    rng = np.random.default_rng(12)
    data = (rng.standard_normal((16, 32, 32)) * 0.9).astype(np.float32)
    case = CASES["gram_charlier_mild"]
    got = run(hip_ctx, data, 0.81, case, arma_amd.FP32)
    ref = oracle(data.astype(np.float64), 0.81, case)
    assert np.max(np.abs(got - ref)) <= 1e-4


def test_nit_tight_tolerance(hip_ctx):
    """eps 1e-12 and 200 steps: both sides converge to the root; the
    results agree to the root's conditioning."""
    rng = np.random.default_rng(13)
    data = rng.standard_normal((4, 16, 16))
    case = CASES["skew_normal_shifted"]
    got = run(hip_ctx, data, 1.0, case, arma_amd.FP64, eps=1e-12, max_iter=200)
    ref = oracle(data, 1.0, case, eps=1e-12, max_iter=200)
    np.testing.assert_allclose(got, ref, atol=1e-9)


def test_nit_rejects_bad_parameters(hip_ctx):
    z = torch.zeros((2, 2, 2), dtype=torch.float64, device=DEV)
    for kw in (dict(interval=(1.0, 1.0)), dict(absolute_error=0.0), dict(max_iterations=0)):
        d = arma_amd.nit_desc(arma_amd.NIT_GRAM_CHARLIER, **kw)
        with pytest.raises(ValueError):
            hip_ctx.nit_realisation(z, (2, 2, 2), arma_amd.FP64, 1.0, d)
