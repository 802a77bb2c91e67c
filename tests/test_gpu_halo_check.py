"""Checked launches of the wavefront recurrence (ARMA_HIP_CHECK_HALO=1).

Tiles exchange their boundary values through per-tile mailboxes in device
memory and publish how many steps' rows are complete; a consumer issues its
halo loads only after it has seen the producer's count.  In a checked launch
every mailbox is first filled with a signalling-NaN bit pattern that no
arithmetic produces, and every consumer checks each halo value it takes for a
slice inside the grid: a value still equal to the pattern was read before its
producer stored it, and the call fails.  Each configuration runs once checked,
at its full size, and must pass and give the same bits as the unchecked
launch (reference: the 7-predecessor readiness of
src/generator/ar_model_parallel.cc:158-171 that the mailboxes replace).
"""
import math
import os

import numpy as np
import pytest
import torch

import arma_amd
from arma_amd import _abi
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def surface_pair(ctx, shape, kind, beta, order, part, seed, precision):
    acf = O.acf(kind, (10, 10, 10), velocity=0.8, beta=beta)
    phi, _ = O.yw_choi(acf, order)
    varwn = O.white_noise_variance(acf, phi)
    n = math.prod(math.ceil(shape[d] / part[d]) for d in range(3))
    params = _abi.dcmt_p521_params(n)
    desc = ctx.ar_desc(shape, phi, varwn, params, seeds=np.full(n, seed, np.uint32),
                       noise_mode=arma_amd.NOISE_PARTITIONED, partition=part,
                       precision=precision, algorithm=arma_amd.AR_WAVEFRONT)
    dt = torch.float64 if precision == arma_amd.FP64 else torch.float32
    plain = torch.empty(shape, dtype=dt, device=DEV)
    ctx.ar_generate(desc, plain)
    checked = torch.empty(shape, dtype=dt, device=DEV)
    os.environ["ARMA_HIP_CHECK_HALO"] = "1"
    try:
        ctx.ar_generate(desc, checked)
        torch.cuda.synchronize()
    finally:
        del os.environ["ARMA_HIP_CHECK_HALO"]
    print(f"checked {shape} phi {phi.shape} precision {precision}: passed", flush=True)
    return plain, checked


@pytest.mark.parametrize("case", [
    # the FP64 benchmark configuration: propagating ACF, phi (7, 7, 2)
    ((1024, 256, 256), O.PROPAGATING, (-0.8, 0.0), 7, (16, 32, 32), 0, arma_amd.FP64),
    # the FP32 precision configuration, as configured and with the
    # standing-wave ACF (7 x 7 x 7 after the fit: the one-lane tile)
    ((2048, 512, 512), O.PROPAGATING, (-0.8, 0.0), 7, (32, 64, 64), 1, arma_amd.FP32),
    ((2048, 512, 512), O.STANDING, (0.8, 0.8), 7, (32, 64, 64), 1, arma_amd.FP32),
    # the CPU-runnable configuration's model in FP64
    ((128, 64, 64), O.STANDING, (0.8, 0.8), 7, (128, 64, 64), 0, arma_amd.FP64),
], ids=["ar-1024x256x256-fp64", "ar-2048x512x512-fp32", "ar-2048x512x512-fp32-standing",
        "ar-128x64x64-fp64-standing"])
def test_checked_launch(hip_ctx, case):
    shape, kind, beta, order, part, seed, precision = case
    plain, checked = surface_pair(hip_ctx, shape, kind, beta, order, part, seed, precision)
    assert torch.equal(plain.view(-1).view(torch.int64 if plain.dtype == torch.float64
                                           else torch.int32),
                       checked.view(-1).view(torch.int64 if checked.dtype == torch.float64
                                             else torch.int32))
    del plain, checked
    torch.cuda.empty_cache()
