"""GPU checks at the full 1024x256x256 size, through properties that do not
depend on the size (the oracle would take minutes there).

On the standing-wave model with its Choi phi (7x7x7):

* AR recurrence (wavefront, FP64): the inverse filter
  eps = zeta - sum_{(k,i,j) != 0} phi(k,i,j) zeta(t-k, x-i, y-j), with zeta
  zero outside the grid, recovers the input noise to 1e-12 of max |eps|.
  The filter is evaluated with torch on the GPU, tap by tap.
* Linear potential (FP64 and FP32): phi(2 zeta) = 2 phi(zeta) to 2^-50 of
  max |phi|.  Multiplying by 2 is exact in every step of the solver: the time
  derivative, the transforms and the window product are linear with fixed
  coefficients, the edge fix-up divides by +-1, and the factor_waves scan
  sees the same event periods (the quadratic's vertex -b/2a is unchanged when
  a and b are doubled), so the wave-number range and hence the window are
  identical.  Only products that fall into the subnormal range lose the
  exactness, far below the bound.
"""
import numpy as np
import pytest
import torch

import arma_amd
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
SHAPE = (1024, 256, 256)
TWO_PI = 2 * np.pi


@pytest.fixture(scope="module")
def surface(hip_ctx):
    # the standing-wave model (stable: |zeta| stays O(1) over the 1024
    # slices; the propagating model of the default workload grows to ~1e80,
    # which no FP32 copy could hold)
    acf = O.acf(O.STANDING, (10, 10, 10), velocity=0.8, beta=(0.8, 0.8))
    phi, _ = O.yw_choi(acf, 7)
    varwn = O.white_noise_variance(acf, phi)
    gen = torch.Generator(device=DEV)
    gen.manual_seed(1234)
    eps = torch.randn(SHAPE, dtype=torch.float64, device=DEV, generator=gen) * np.sqrt(varwn)
    z = eps.clone()
    d = hip_ctx.ar_desc(SHAPE, phi, varwn, [arma_amd.mt19937_params()],
                        algorithm=arma_amd.AR_WAVEFRONT)
    print(f"full-size AR {SHAPE} phi {phi.shape}", flush=True)
    hip_ctx.ar_recurrence(d, z)
    torch.cuda.synchronize()
    return eps, z, phi


def test_ar_inverse_filter_full_size(surface):
    eps, z, phi = surface
    res = z.clone()
    nt, nx, ny = SHAPE
    for (k, i, j), c in np.ndenumerate(phi):
        if (k, i, j) == (0, 0, 0) or c == 0.0:
            continue
        res[k:, i:, j:] -= float(c) * z[:nt - k, :nx - i, :ny - j]
    err = float((res - eps).abs().max() / eps.abs().max())
    assert err <= 1e-12, err


@pytest.mark.parametrize("precision", [arma_amd.FP64, arma_amd.FP32])
def test_linear_potential_homogeneous_full_size(hip_ctx, surface, precision):
    _, z, _ = surface
    dt = torch.float64 if precision == arma_amd.FP64 else torch.float32
    zeta = z.to(dt)
    npts = (SHAPE[0], 16)
    d = hip_ctx.pot_desc(SHAPE, (TWO_PI,) * 3, 12.0, (0.0, -10.0), (SHAPE[0] - 1.0, 0.0), npts,
                         precision)
    out1 = torch.empty((npts[0], npts[1], SHAPE[1], SHAPE[2]), dtype=dt, device=DEV)
    out2 = torch.empty_like(out1)
    print(f"full-size potential {out1.shape} {dt}", flush=True)
    wn1 = hip_ctx.linear_potential(d, zeta, out1, want_wnmax=True)
    wn2 = hip_ctx.linear_potential(d, zeta * 2, out2, want_wnmax=True)
    np.testing.assert_array_equal(wn1, wn2)
    scale = float(out1.abs().max())
    assert scale > 0 and np.isfinite(scale)
    diff = float((out2 - 2 * out1).abs().max())
    assert diff <= 2.0 ** -50 * 2 * scale, (diff, scale)
