"""Pins the oracle's ACF_generator (src/generator/acf_generator.cc:130-215):
the ACF as the auto-covariance of a synthetic plain wave with exponential
decay.  Checks: a numpy restatement (numpy.fft, ddof=1 variance) for every
wave profile and both shape modes, and the circular auto-covariance by brute
force on a tiny grid.
"""
import numpy as np
import pytest

import pyoracle as O


def wave_np(func, amp, vel, k, r, ws):
    t, x, y = (np.linspace(-r[q], r[q], ws[q] + 1) if ws[q] else np.full(1, -r[q])
               for q in range(3))
    d = [2 * r[q] / ws[q] for q in range(3)]
    t = -r[0] + d[0] * np.arange(ws[0] + 1)
    x = -r[1] + d[1] * np.arange(ws[1] + 1)
    y = -r[2] + d[2] * np.arange(ws[2] + 1)
    T, X, Y = np.meshgrid(t, x, y, indexing="ij")
    th = 2 * np.pi * (k[0] * X + k[1] * Y - vel * T)
    if func == O.WAVE_SIN:
        w = amp * np.sin(th)
    elif func == O.WAVE_COS:
        w = amp * np.sin(th + 0.5 * np.pi)
    elif func == O.WAVE_STOKES:
        st = 2 * np.pi * amp * np.hypot(k[0], k[1])
        w = amp * (np.cos(th) + 0.5 * st * np.cos(2 * th) + 3 / 8 * st * st * np.cos(3 * th))
    else:
        w = amp * np.cos(2 * np.pi * (k[0] * X + k[1] * Y)) * np.cos(2 * np.pi * vel * T)
    return w, (T, X, Y)


def acfgen_np(func, amp, vel, k, alpha, nwaves, shape):
    if all(s > 0 for s in shape):
        ws = [2 * s for s in shape]
        w, txy = wave_np(func, amp, vel, k, nwaves, ws)
    else:
        ws, var0, var = [2, 2, 2], -1.0, -1.0
        while all(v < 64 for v in ws):
            w, txy = wave_np(func, amp, vel, k, nwaves, ws)
            var0, var = var, np.var(w, ddof=1)
            if not var0 < 0 and abs(var - var0) < 1e-3:
                break
            ws = [2 * v for v in ws]
    T, X, Y = txy
    we = w * np.exp(-(np.abs(T * alpha[0]) + np.abs(X * alpha[1]) + np.abs(Y * alpha[2])))
    we *= np.sqrt(np.var(w, ddof=1) / np.var(we, ddof=1))
    n = we.size
    acf = np.real(np.fft.ifftn(np.abs(np.fft.fftn(we)) ** 2)) * n / n / n
    h = [s // 2 for s in we.shape]
    return acf[:h[0], :h[1], :h[2]], we


@pytest.mark.parametrize("func", [O.WAVE_SIN, O.WAVE_COS, O.WAVE_STOKES, O.WAVE_STANDING])
def test_acf_generator_against_numpy(func):
    args = dict(amplitude=3.0, velocity=0.5, wavenum=(0.0625, 0.3), alpha=(2.0, 0.2, 1.0),
                nwaves=(1.85, 4.0, 1.0), shape=(6, 8, 5))
    got = O.acf_generate(func, **args)
    ref, _ = acfgen_np(func, 3.0, 0.5, (0.0625, 0.3), (2.0, 0.2, 1.0), (1.85, 4.0, 1.0), (6, 8, 5))
    assert got.shape == (6, 8, 5)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_acf_generator_optimal_size():
    """no shape: the wave size doubles until its variance settles."""
    got = O.acf_generate(O.WAVE_COS, shape=(0, 0, 0))
    ref, _ = acfgen_np(O.WAVE_COS, 1.0, 1.0, (0.8, 0.0), (0.06,) * 3, (1.5,) * 3, (0, 0, 0))
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_auto_covariance_brute_force():
    """acf(tau) = (1/n) sum_x w(x) w(x + tau) (circular), and acf(0) is the
    mean square of the decayed wave."""
    shape = (2, 3, 2)
    got = O.acf_generate(O.WAVE_STOKES, amplitude=0.7, velocity=0.3, wavenum=(0.2, 0.1),
                         alpha=(0.3, 0.1, 0.2), nwaves=(1.0, 1.2, 0.8), shape=shape)
    _, we = acfgen_np(O.WAVE_STOKES, 0.7, 0.3, (0.2, 0.1), (0.3, 0.1, 0.2), (1.0, 1.2, 0.8),
                      shape)
    s = we.shape
    for a in range(shape[0]):
        for b in range(shape[1]):
            for c in range(shape[2]):
                acc = 0.0
                for i in range(s[0]):
                    for j in range(s[1]):
                        for k in range(s[2]):
                            acc += we[i, j, k] * we[(i + a) % s[0], (j + b) % s[1], (k + c) % s[2]]
                assert got[a, b, c] == pytest.approx(acc / we.size, rel=1e-12, abs=1e-14)
    assert got[0, 0, 0] == pytest.approx(np.mean(we ** 2), rel=1e-12)


# ---- the library's host implementation against the oracle (CPU only) ----

@pytest.mark.parametrize("func", [0, 1, 2, 3])
def test_library_acf_generate_matches_oracle(func):
    from arma_amd import _abi
    args = dict(amplitude=3.0, velocity=0.5, alpha=(2.0, 0.2, 1.0), nwaves=(1.85, 4.0, 1.0),
                shape=(6, 8, 5))
    got, delta = _abi.acf_generate(func, beta=(0.0625, 0.3), **args)
    ref = O.acf_generate(func, wavenum=(0.0625, 0.3), **args)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_allclose(delta, [1.85 / 5, 4.0 / 7, 1.0 / 4], rtol=1e-15)


def test_library_acf_generate_defaults_and_doubling():
    from arma_amd import _abi
    got, _ = _abi.acf_generate(shape=(0, 0, 0))
    ref = O.acf_generate(O.WAVE_COS, shape=(0, 0, 0))
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    got, _ = _abi.acf_generate()
    assert got.shape == (16, 16, 16)
    with pytest.raises(ValueError):
        _abi.acf_generate(func=7)


def test_reference_acf_generator_case():
    """src/test/acf-generator-test.cc's input (it checks only that the
    generator runs): cos, amplitude 4, alpha (2,1,10), velocity 1,
    beta (0.4,0), nwaves 12.5, shape (32,1,1)."""
    from arma_amd import _abi
    args = dict(amplitude=4.0, velocity=1.0, alpha=(2.0, 1.0, 10.0), nwaves=(12.5,) * 3,
                shape=(32, 1, 1))
    got, delta = _abi.acf_generate(_abi.WAVE_COS, beta=(0.4, 0.0), **args)
    ref = O.acf_generate(O.WAVE_COS, wavenum=(0.4, 0.0), **args)
    assert got.shape == (32, 1, 1) and np.all(np.isfinite(got))
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    assert delta[0] == pytest.approx(12.5 / 31)


# ---- the GPU auto-covariance (arma_hip_acf_generate_device) ----

@pytest.mark.gpu
@pytest.mark.parametrize("func,shape", [(0, (6, 8, 5)), (1, (32, 1, 1)), (2, (3, 20, 7)),
                                        (3, (32, 32, 32)), (1, (0, 0, 0))])
def test_gpu_acf_generate_matches_oracle(hip_ctx, func, shape):
    """The direct circular sums on the GPU against the oracle's FFT-based
    auto-covariance: every wave profile, odd extents (2 shape + 1, up to 65
    per axis and lag tiles with lanes past the crop), and the doubling mode."""
    from arma_amd import _abi
    args = dict(amplitude=3.0, velocity=0.5, alpha=(2.0, 0.2, 1.0), nwaves=(1.85, 4.0, 1.0),
                shape=shape)
    got, delta = _abi.acf_generate(func, beta=(0.0625, 0.3), ctx=hip_ctx, **args)
    host, hdelta = _abi.acf_generate(func, beta=(0.0625, 0.3), **args)
    ref = O.acf_generate(func, wavenum=(0.0625, 0.3), **args)
    assert got.shape == ref.shape and delta == hdelta
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_allclose(got, host, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
