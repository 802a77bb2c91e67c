"""GPU parity of the linear and high-amplitude velocity potentials, through
the C ABI.

FP64 bar: max |phi_gpu - phi_oracle| <= 1e-10 * max |phi_oracle| (the GPU
transform is a half-spectrum C2R of the symmetrised window; the oracle does
the reference's full complex transforms).  FP32 bar: 2e-4 relative to max.
The per-slice wave-number range from factor_waves must agree to 1e-15.
"""
import numpy as np
import pytest
import torch

import arma_amd
import pyoracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TWO_PI = 2 * np.pi


def ar_surface(shape, seed):
    rng = np.random.default_rng(seed)
    acf = O.acf(O.PROPAGATING, (10, 10, 10), velocity=0.8, beta=(-0.8, 0.0))
    phi, _ = O.yw_choi(acf, 4)
    return O.ar_generate(rng.standard_normal(shape), phi)


def run_gpu(ctx, zeta, depth, lb, ub, npts, precision=arma_amd.FP64, want_wn=False,
            solver="linear", delta=(TWO_PI,) * 3):
    dt = torch.float64 if precision == arma_amd.FP64 else torch.float32
    z = torch.from_numpy(np.ascontiguousarray(zeta)).to(DEV, dt)
    out = torch.empty((npts[0], npts[1], zeta.shape[1], zeta.shape[2]), dtype=dt, device=DEV)
    d = ctx.pot_desc(zeta.shape, delta, depth, lb, ub, npts, precision)
    wn = ctx.linear_potential(d, z, out, want_wnmax=want_wn, solver=solver)
    return out.cpu().numpy().astype(np.float64), wn


@pytest.mark.parametrize("shape,npts,lb,ub", [
    ((6, 16, 16), (6, 4), (0.0, -10.0), (5.0, 0.0)),
    ((9, 32, 64), (3, 5), (2.0, -8.0), (6.0, 1.0)),
    ((2, 8, 16), (2, 3), (0.0, -2.0), (1.0, 0.0)),
    ((12, 64, 32), (12, 16), (0.0, -10.0), (11.0, 0.0)),
])
def test_potential_fp64_matches_oracle(hip_ctx, shape, npts, lb, ub):
    zeta = ar_surface(shape, 1)
    got, wn = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, want_wn=True)
    ref, wref = O.linear_potential(zeta, (TWO_PI,) * 3, 12.0, lb, ub, npts, return_wnmax=True)
    np.testing.assert_allclose(wn, wref, rtol=1e-15)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-10 * scale
    # the fixed-up edges (row 0, column 0) on their own
    assert np.max(np.abs(got[:, :, 0, :] - ref[:, :, 0, :])) <= 1e-10 * scale
    assert np.max(np.abs(got[:, :, :, 0] - ref[:, :, :, 0])) <= 1e-10 * scale


@pytest.mark.parametrize("shape,nz", [
    ((2, 1024, 16), 20),  # four elements per lane: every level group seeds the window
    ((3, 512, 32), 20),   # two per lane: the window runs on across groups of 8 levels
    ((4, 64, 32), 40),    # seeds at levels 0, 16 and 32
    ((2, 64, 256), 3),    # edge fix-up over two row workgroups of 32 rows
    ((2, 16, 1024), 3),   # rows of 1024 (32-point first DFTs in the row pass)
])
@pytest.mark.parametrize("solver", ["linear", "high_amplitude"])
def test_potential_long_transforms_and_many_levels(hip_ctx, shape, nz, solver):
    print(f"potential {solver} {shape} nz={nz}", flush=True)
    zeta = ar_surface(shape, 3)
    npts, lb, ub = (shape[0], nz), (0.0, -10.0), (float(shape[0] - 1), 0.0)
    delta = (TWO_PI, 0.5, 0.5)
    got, _ = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, solver=solver, delta=delta)
    fn = O.linear_potential if solver == "linear" else O.high_amplitude_potential
    ref = fn(zeta, delta, 12.0, lb, ub, npts)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_potential_fp32(hip_ctx):
    zeta = ar_surface((8, 64, 64), 2)
    npts, lb, ub = (8, 8), (0.0, -10.0), (7.0, 0.0)
    got, _ = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, precision=arma_amd.FP32)
    ref = O.linear_potential(zeta, (TWO_PI,) * 3, 12.0, lb, ub, npts)
    err = np.abs(got - ref)
    assert np.max(err) <= 2e-4 * np.max(np.abs(ref))
    assert np.sqrt(np.mean(err ** 2)) <= 1e-5 * np.max(np.abs(ref))


def test_potential_single_mode(hip_ctx):
    nt, nx, ny, p, q = 5, 32, 16, 3, 2
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    mode = np.cos(2 * np.pi * (p * x / nx + q * y / ny)) + 0.3 * np.sin(2 * np.pi * x / nx)
    zeta = (np.arange(nt)[:, None, None] + 1.0) * mode[None]
    got, wn = run_gpu(hip_ctx, zeta, 12.0, (2.0, -4.0), (2.0, -1.0), (1, 2), want_wn=True)
    for j, z in enumerate([-4.0, -1.0]):
        w = O.window(nx, ny, wn[0, 0], wn[0, 1], z, 12.0)

        def ws(a, b):
            return 0.5 * (w[a % nx, b % ny] + w[-a % nx, -b % ny])
        expect = -(np.cos(2 * np.pi * (p * x / nx + q * y / ny)) * ws(p, q) +
                   0.3 * np.sin(2 * np.pi * x / nx) * ws(1, 0)) / TWO_PI
        np.testing.assert_allclose(got[0, j][1:, 1:], expect[1:, 1:], rtol=1e-10, atol=1e-11)


def test_potential_host_pointers(hip_ctx):
    zeta = ar_surface((4, 16, 16), 3)
    npts = (4, 3)
    d = hip_ctx.pot_desc(zeta.shape, (TWO_PI,) * 3, 12.0, (0.0, -5.0), (3.0, 0.0), npts)
    out = np.empty((4, 3, 16, 16), np.float64)
    hip_ctx.linear_potential(d, np.ascontiguousarray(zeta), out, on_device=False)
    ref = O.linear_potential(zeta, (TWO_PI,) * 3, 12.0, (0.0, -5.0), (3.0, 0.0), npts)
    assert np.max(np.abs(out - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("precision", [arma_amd.FP64, arma_amd.FP32])
def test_potential_unaligned_surface(hip_ctx, precision):
    """A surface that starts at an odd element of its buffer (as in a batch
    arena): the forward transform reads it element by element."""
    dt = torch.float64 if precision == arma_amd.FP64 else torch.float32
    zeta = ar_surface((4, 16, 16), 8)
    npts = (4, 3)
    buf = torch.zeros(zeta.size + 1, dtype=dt, device=DEV)
    z = buf[1:].view(zeta.shape)
    z.copy_(torch.from_numpy(zeta))
    assert z.data_ptr() % (2 * z.element_size()) != 0
    out = torch.empty((4, 3, 16, 16), dtype=dt, device=DEV)
    d = hip_ctx.pot_desc(zeta.shape, (TWO_PI,) * 3, 12.0, (0.0, -5.0), (3.0, 0.0), npts, precision)
    hip_ctx.linear_potential(d, z, out)
    aligned, _ = run_gpu(hip_ctx, z.cpu().numpy().astype(np.float64), 12.0, (0.0, -5.0),
                         (3.0, 0.0), npts, precision=precision)
    # the same values as from an aligned copy, bit for bit
    np.testing.assert_array_equal(out.cpu().numpy().astype(np.float64), aligned)


def test_potential_errors(hip_ctx):
    zeta = ar_surface((4, 16, 16), 4)
    with pytest.raises(ValueError):
        run_gpu(hip_ctx, zeta, 5.0, (0.0, -6.0), (3.0, 0.0), (4, 2))
    with pytest.raises(ValueError):
        run_gpu(hip_ctx, ar_surface((4, 12, 16), 4), 12.0, (0.0, -6.0), (3.0, 0.0), (4, 2))


# ---- high-amplitude solver ----

@pytest.mark.parametrize("shape,npts,lb,ub,delta", [
    ((6, 16, 16), (6, 4), (0.0, -10.0), (5.0, 0.0), (TWO_PI,) * 3),
    ((9, 32, 64), (3, 5), (2.0, -8.0), (6.0, 1.0), (TWO_PI, 0.5, 0.5)),
    ((2, 8, 16), (2, 3), (0.0, -2.0), (1.0, 0.0), (1.0, 0.3, 0.7)),
    ((12, 64, 32), (12, 16), (0.0, -10.0), (11.0, 0.0), (TWO_PI, 1.0, 0.25)),
])
def test_high_amplitude_fp64_matches_oracle(hip_ctx, shape, npts, lb, ub, delta):
    print(f"high_amplitude fp64 {shape} {npts}", flush=True)
    zeta = ar_surface(shape, 5)
    got, wn = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, want_wn=True, solver="high_amplitude",
                      delta=delta)
    ref, wref = O.high_amplitude_potential(zeta, delta, 12.0, lb, ub, npts, return_wnmax=True)
    np.testing.assert_allclose(wn, wref, rtol=1e-15)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-10 * scale
    # the fixed-up edges (row 0, column 0) on their own
    assert np.max(np.abs(got[:, :, 0, :] - ref[:, :, 0, :])) <= 1e-10 * scale
    assert np.max(np.abs(got[:, :, :, 0] - ref[:, :, :, 0])) <= 1e-10 * scale


def test_high_amplitude_fp32(hip_ctx):
    zeta = ar_surface((8, 64, 64), 6)
    npts, lb, ub, delta = (8, 8), (0.0, -10.0), (7.0, 0.0), (TWO_PI, 0.5, 0.5)
    got, _ = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, precision=arma_amd.FP32,
                     solver="high_amplitude", delta=delta)
    ref = O.high_amplitude_potential(zeta, delta, 12.0, lb, ub, npts)
    err = np.abs(got - ref)
    assert np.max(err) <= 2e-4 * np.max(np.abs(ref))
    assert np.sqrt(np.mean(err ** 2)) <= 1e-5 * np.max(np.abs(ref))


@pytest.mark.parametrize("a,b,c", [(0.3, -0.2, 0.7), (1.5, 2.0, -0.4)])
def test_high_amplitude_plane(hip_ctx, a, b, c):
    """Plane zeta = a x + b y + c t: constant multiplier, phi = 4 pi Re(c / D)."""
    delta = (TWO_PI, 0.5, 0.25)
    nt, nx, ny = 5, 8, 16
    t = np.arange(nt)[:, None, None] * delta[0]
    x = np.arange(nx)[None, :, None] * delta[1]
    y = np.arange(ny)[None, None, :] * delta[2]
    zeta = a * x + b * y + c * t + 0.0 * (t + x + y)
    got, _ = run_gpu(hip_ctx, zeta, 12.0, (0.0, -6.0), (4.0, 0.0), (5, 3),
                     solver="high_amplitude", delta=delta)
    s = np.sqrt(1 + a * a + b * b)
    expect = 4 * np.pi * (c / complex(-1 / s, (a + b) / s - a - b)).real
    np.testing.assert_allclose(got, expect, rtol=1e-12)


def test_high_amplitude_host_pointers_and_errors(hip_ctx):
    zeta = ar_surface((4, 16, 16), 7)
    npts = (4, 3)
    d = hip_ctx.pot_desc(zeta.shape, (TWO_PI,) * 3, 12.0, (0.0, -5.0), (3.0, 0.0), npts)
    out = np.empty((4, 3, 16, 16), np.float64)
    hip_ctx.high_amplitude_potential(d, np.ascontiguousarray(zeta), out, on_device=False)
    ref = O.high_amplitude_potential(zeta, (TWO_PI,) * 3, 12.0, (0.0, -5.0), (3.0, 0.0), npts)
    assert np.max(np.abs(out - ref)) <= 1e-10 * np.max(np.abs(ref))
    with pytest.raises(ValueError):
        run_gpu(hip_ctx, zeta, 5.0, (0.0, -6.0), (3.0, 0.0), (4, 2), solver="high_amplitude")


@pytest.mark.parametrize("solver", ["linear", "high_amplitude"])
def test_potential_window_overflow(hip_ctx, solver):
    """z up to 5 with wave numbers up to ~270: the window overflows at the
    high wave numbers of the upper levels, where the reference zeroes it
    (src/velocity/linear_solver.cc:176-183).  The surface is scaled by
    1e-250 so that the surviving huge window values leave a finite field;
    every (t, z) plane must match the oracle relative to its own maximum."""
    zeta = ar_surface((6, 16, 16), 1) * 1e-250
    npts, lb, ub = (6, 7), (0.0, -10.0), (5.0, 5.0)
    got, _ = run_gpu(hip_ctx, zeta, 12.0, lb, ub, npts, solver=solver)
    ref = O.linear_potential(zeta, (TWO_PI,) * 3, 12.0, lb, ub, npts, solver=solver)
    assert np.isfinite(ref).all() and np.isfinite(got).all()
    for i in range(npts[0]):
        for j in range(npts[1]):
            scale = np.max(np.abs(ref[i, j]))
            assert np.max(np.abs(got[i, j] - ref[i, j])) <= 1e-9 * scale, (i, j)
