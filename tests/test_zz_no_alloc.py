"""arma_hip_ar_generate_noalloc against arma_hip_ar_generate, through the C ABI.

The no-allocation entry point runs the same noise and wavefront kernels as
arma_hip_ar_generate with every working buffer carved from the library's
static device scratch, so its surfaces are bit-identical.  Two calls in a row
with the second smaller check that what the first call left in the shared
scratch does not leak into the second (the progress counters and the zero
word are cleared at each call).
"""
import math

import numpy as np
import pytest
import torch

import arma_amd
from arma_amd import _abi
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def desc_for(ctx, shape, order, part, seed, precision):
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, order)
    varwn = O.white_noise_variance(acf, phi)
    n = math.prod(math.ceil(shape[d] / part[d]) for d in range(3))
    params = _abi.dcmt_p521_params(n)
    return ctx.ar_desc(shape, phi, varwn, params, seeds=np.arange(n, dtype=np.uint32) + seed,
                       noise_mode=arma_amd.NOISE_PARTITIONED, partition=part,
                       precision=precision, algorithm=arma_amd.AR_WAVEFRONT)


@pytest.mark.parametrize("precision", [arma_amd.FP64, arma_amd.FP32])
def test_noalloc_equals_generate(hip_ctx, precision):
    dt = torch.float64 if precision == arma_amd.FP64 else torch.float32
    # the second call is smaller than the first
    for shape, order, part, seed in [((96, 160, 144), 7, (32, 40, 48), 11),
                                     ((40, 48, 33), 5, (20, 16, 11), 3)]:
        print(f"noalloc {shape} order {order} precision {precision}", flush=True)
        d = desc_for(hip_ctx, shape, order, part, seed, precision)
        want = torch.empty(shape, dtype=dt, device=DEV)
        hip_ctx.ar_generate(d, want)
        got = torch.full(shape, float("nan"), dtype=dt, device=DEV)
        hip_ctx.ar_generate_noalloc(d, got)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())


def test_noalloc_rejects_oversized_surface(hip_ctx):
    # 512^3 FP64 at order 7 needs about 700 MB of halo mailboxes: refused on
    # the host before anything is enqueued, the (full-size) output untouched
    shape = (512, 512, 512)
    d = desc_for(hip_ctx, shape, 7, (512, 512, 512), 0, arma_amd.FP64)
    z = torch.zeros(shape, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError, match="no-allocation entry point has"):
        hip_ctx.ar_generate_noalloc(d, z)
    assert torch.count_nonzero(z).item() == 0


def test_noalloc_rejects_plain_algorithm(hip_ctx):
    d = desc_for(hip_ctx, (16, 16, 16), 3, (16, 16, 16), 0, arma_amd.FP64)
    d.algorithm = arma_amd.AR_PLAIN
    z = torch.zeros((16, 16, 16), dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError, match="wavefront"):
        hip_ctx.ar_generate_noalloc(d, z)
