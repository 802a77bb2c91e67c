"""Pins the oracle's linear velocity-potential solver.

The reference's Linear_solver (src/velocity/linear_solver.cc:20-184) has no
tests of its own; these checks pin the transform convention (numpy golden
vectors), the window (closed form cosh(l(z+h))/cosh(lh) and the rotate map
of src/blitz.hh:54-71), the derivative (exact on quadratics), factor_waves
(a run-by-run restatement of src/factor_waves.cc:8-89 on random slices; the
hand-worked known answers are in test_oracle_pins.py) and the whole solver
(a single Fourier mode has an analytic answer; numpy restatements with the
edge fix-up in closed form).
"""
import os

import numpy as np
import pytest

import pyoracle as O


@pytest.fixture(scope="module")
def fftgold(golden_dir):
    return np.load(os.path.join(golden_dir, "numpy_fft.npz"))


@pytest.mark.parametrize("key", ["16x16", "8x12", "5x7"])
def test_fft_against_numpy(fftgold, key):
    x = fftgold[f"in_{key}"]
    np.testing.assert_allclose(O.fft2(x, -1), fftgold[f"fwd_{key}"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.fft2(x, +1), fftgold[f"bwd_{key}"], rtol=1e-12, atol=1e-12)


def rotate_ref(a):
    """src/blitz.hh:54-71 written as an index map."""
    nx, ny = a.shape
    mr, mc = nx // 2, ny // 2
    b = a.copy()
    for i in range(nx):
        for j in range(mc):
            b[i, j], b[i, (j + mc) % ny] = b[i, (j + mc) % ny], b[i, j]
    for i in range(mr):
        for j in range(ny):
            b[i, j], b[(i + mr) % nx, j] = b[(i + mr) % nx, j], b[i, j]
    return b


@pytest.mark.parametrize("nx,ny", [(16, 16), (8, 12)])
def test_window_closed_form(nx, ny):
    wx, wy, z, h = 0.31, 0.27, -3.5, 12.0
    got = O.window(nx, ny, wx, wy, z, h)
    u = np.arange(nx)[:, None] * (wx / (nx - 1))
    v = np.arange(ny)[None, :] * (wy / (ny - 1))
    l = 2 * np.pi * np.sqrt(u * u + v * v)
    w = 4 * np.pi * np.cosh(l * (z + h)) / np.cosh(l * h)
    w[0, 0] = w[1, 1]          # W(0,0) := W(1,1)
    np.testing.assert_allclose(got, np.roll(np.roll(w, ny // 2, axis=1), nx // 2, axis=0), rtol=1e-12)
    np.testing.assert_allclose(got, rotate_ref(w), rtol=1e-12)


def test_rotate_odd_is_half_swap():
    # for odd extents blitz::rotate swaps halves and leaves the last line
    a = np.arange(15.0).reshape(3, 5)
    b = rotate_ref(a)
    assert list(b[:, 4]) == [9.0, 4.0, 14.0]   # last column only row-rotated
    w = O.window(3, 5, 0.1, 0.1, -1.0, 5.0)
    assert w.shape == (3, 5)


def test_derivative_exact_on_quadratic():
    nt, nx, ny, dt = 6, 4, 3, 2 * np.pi
    t = np.arange(nt)[:, None, None] * dt
    a = np.random.default_rng(0).standard_normal((1, nx, ny))
    f = a * t ** 2 + 3.0 * t + 1.0
    for it in range(nt):
        np.testing.assert_allclose(O.derivative_t(f, dt, it), (2 * a * t[it] + 3.0)[0], rtol=1e-10)
    # fewer than 3 slices: the slice itself
    np.testing.assert_array_equal(O.derivative_t(f[:2], dt, 1), f[1])


def factor_waves_ref(z, t, dt):
    """Independent restatement of src/factor_waves.cc:103-158 (one axis)."""
    g = 9.82
    events = []
    nx, ny = z.shape
    for y in range(ny):
        for x in range(1, nx - 1):
            z0, z1, z2 = z[x, y], z[x - 1, y], z[x + 1, y]
            dw1, dw2 = z0 - z1, z0 - z2
            if (dw1 > 0) == (dw2 > 0):
                a = -0.5 * (dw1 + dw2) / (dt * dt)
                b = dw1 / dt - a * dt * (2.0 * t - 1.0)
                c = z0 - t * dt * (a * t * dt + b)
                tex = -0.5 * b / a
                events.append((tex, c + tex * (b + a * tex)))
    # runs of same-sign elevation; representative = first max |elevation|
    runs = []
    for tex, wex in events:
        if runs and (runs[-1][0] == (wex > 0)):
            if abs(runs[-1][2]) < abs(wex):
                runs[-1] = (runs[-1][0], tex, wex)
        else:
            runs.append((wex > 0, tex, wex))
    ks = []
    # a pair (r, r-1) is scored when run r is closed by a sign change
    for r in range(1, len(runs) - 1):
        w = 2 * np.pi / abs((runs[r][1] - runs[r - 1][1]) * 2.0)
        ks.append(w * w / g)
    return (min(ks) if ks else np.finfo(float).max), (max(ks) if ks else np.finfo(float).tiny)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_factor_waves_restatement(seed):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((12, 9))
    t, dt = 37, 2 * np.pi
    got = O.factor_waves(z, t, dt)
    kx = factor_waves_ref(z, t, dt)
    ky = factor_waves_ref(z.T, t, dt)
    np.testing.assert_allclose(got, (kx[0], ky[0], kx[1], ky[1]), rtol=1e-13)


def edges_numpy(res):
    """Edge fix-up of src/velocity/basic_solver.cc:61-68 in the closed forms
    that tests/test_oracle_pins.py checks against hand-worked values."""
    r = res.copy()
    r[0, 0] = 3 * r[1, 1] - r[2, 1] - r[1, 2]
    for k in range(1, r.shape[0]):
        r[k, 0] = r[k, 1] + r[k - 1, 1] - r[k - 1, 2]
    for l in range(1, r.shape[1]):
        r[0, l] = r[1, l] + r[1, l - 1] - r[2, l - 1]
    return r


def potential_numpy(zeta, delta, depth, lb, ub, npts):
    """numpy restatement of the linear solver (independent of the oracle)."""
    nt, nx, ny = zeta.shape
    dtd = (ub[0] - lb[0]) / max(npts[0] - 1, 1)
    dzd = (ub[1] - lb[1]) / max(npts[1] - 1, 1)
    out = np.empty((npts[0], npts[1], nx, ny))
    for i in range(npts[0]):
        it = int(lb[0] + dtd * i)
        if nt < 3:
            d = zeta[it]
        elif it == 0:
            d = 0.5 * (-zeta[2] + 4 * zeta[1] - 3 * zeta[0]) / delta[0]
        elif it == nt - 1:
            d = 0.5 * (3 * zeta[it] - 4 * zeta[it - 1] + zeta[it - 2]) / delta[0]
        else:
            d = 0.5 * (zeta[it + 1] - zeta[it - 1]) / delta[0]
        kx = factor_waves_ref(zeta[it], int(delta[0] * it), delta[0])[0]
        ky = factor_waves_ref(zeta[it].T, int(delta[0] * it), delta[0])[0]
        F = np.fft.fft2(-d)
        u = np.arange(nx)[:, None] * ((1 / kx) / (nx - 1))
        v = np.arange(ny)[None, :] * ((1 / ky) / (ny - 1))
        l = 2 * np.pi * np.sqrt(u * u + v * v)
        for j in range(npts[1]):
            z = lb[1] + dzd * j
            with np.errstate(over="ignore", invalid="ignore"):
                # cosh(l(z+h))/cosh(lh) in the overflow-safe exp form of
                # src/velocity/linear_solver.cc:139-150
                e1, e2 = np.exp(-l * z), np.exp(-2.0 * l * depth)
                w = 4 * np.pi * (1.0 + e1 * e1 * e2) / (e1 * (1.0 + e2))
            w[0, 0] = w[1, 1]
            w[~np.isfinite(w)] = 0
            w = rotate_ref(w)
            res = np.real(np.fft.ifft2(F * w))
            out[i, j] = edges_numpy(res)
    return out


def test_linear_potential_against_numpy():
    rng = np.random.default_rng(7)
    zeta = O.ar_generate(rng.standard_normal((6, 16, 16)),
                         np.array([[[0, 0.5], [0.3, -0.2]], [[0.2, 0.1], [0.1, 0.05]]]))
    delta = (2 * np.pi,) * 3
    args = (12.0, (0.0, -10.0), (5.0, 0.0), (6, 4))
    got = O.linear_potential(zeta, delta, *args)
    ref = potential_numpy(zeta, delta, *args)
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


def test_linear_potential_single_mode():
    nt, nx, ny, p, q = 5, 16, 16, 3, 2
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    mode = np.cos(2 * np.pi * (p * x / nx + q * y / ny)) + 0.3 * np.sin(2 * np.pi * (x / nx))
    t = np.arange(nt)[:, None, None]
    zeta = (t + 1.0) * mode[None]
    delta = (2 * np.pi,) * 3
    out, wn = O.linear_potential(zeta, delta, 12.0, (2.0, -4.0), (2.0, -1.0), (1, 2), return_wnmax=True)
    for j, z in enumerate([-4.0, -1.0]):
        w = O.window(nx, ny, wn[0, 0], wn[0, 1], z, 12.0)
        ws = lambda a, b: 0.5 * (w[a % nx, b % ny] + w[-a % nx, -b % ny])
        expect = -(np.cos(2 * np.pi * (p * x / nx + q * y / ny)) * ws(p, q) +
                   0.3 * np.sin(2 * np.pi * x / nx) * ws(1, 0)) / delta[0]
        np.testing.assert_allclose(out[0, j][1:, 1:], expect[1:, 1:], rtol=1e-10, atol=1e-10)


def test_potential_rejects_z_below_depth():
    zeta = np.zeros((3, 8, 8))
    with pytest.raises(ValueError):
        O.linear_potential(zeta, (1.0, 1.0, 1.0), 5.0, (0.0, -6.0), (2.0, 0.0), (3, 2))


# ---- high-amplitude solver (src/velocity/high_amplitude_solver.cc:6-23) ----


def test_derivative_xy_exact_on_quadratic():
    nt, nx, ny, dx, dy = 3, 7, 5, 0.7, 1.3
    x = np.arange(nx)[:, None] * dx
    y = np.arange(ny)[None, :] * dy
    a, b, c = 0.3, -1.1, 0.45
    f = np.broadcast_to(a * x ** 2 + b * y ** 2 + c * x * y + 2.0 * x - y + 0.5, (nt, nx, ny))
    for it in range(nt):
        np.testing.assert_allclose(O.derivative_xy(f, dx, 1, it), 2 * a * x + c * y + 2.0 + 0 * y,
                                   rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(O.derivative_xy(f, dy, 2, it), 2 * b * y + c * x - 1.0 + 0 * x,
                                   rtol=1e-12, atol=1e-12)
    # fewer than 3 points along the dimension: the slice itself
    g = np.random.default_rng(1).standard_normal((2, 2, 6))
    np.testing.assert_array_equal(O.derivative_xy(g, 1.0, 1, 1), g[1])


def plane(nt, nx, ny, a, b, c, delta):
    t = np.arange(nt)[:, None, None] * delta[0]
    x = np.arange(nx)[None, :, None] * delta[1]
    y = np.arange(ny)[None, None, :] * delta[2]
    return a * x + b * y + c * t + 0.0 * (t + x + y)


def ha_plane_value(a, b, c):
    """phi of the plane zeta = a x + b y + c t: the multiplier
    F = c / (i((a + b)/s - a - b) - 1/s) is constant, so only the (0,0)
    bin survives, and with no extrema factor_waves leaves kmin at the
    largest double, the wave numbers vanish and the window is 4 pi."""
    s = np.sqrt(1 + a * a + b * b)
    return 4 * np.pi * (c / complex(-1 / s, (a + b) / s - a - b)).real


@pytest.mark.parametrize("a,b,c", [(0.3, -0.2, 0.7), (1.5, 2.0, -0.4), (-0.05, 0.02, 1.0)])
def test_high_amplitude_plane_closed_form(a, b, c):
    delta = (2 * np.pi, 0.5, 0.25)
    zeta = plane(5, 8, 16, a, b, c, delta)
    args = (12.0, (0.0, -6.0), (4.0, 0.0), (5, 3))
    got = O.high_amplitude_potential(zeta, delta, *args)
    np.testing.assert_allclose(got, ha_plane_value(a, b, c), rtol=1e-12)
    lin = O.linear_potential(zeta, delta, *args)
    np.testing.assert_allclose(lin, -4 * np.pi * c, rtol=1e-12)


def test_high_amplitude_small_slope_limit():
    """F = zeta_t / D with D = -1 + O(slope^2): the high-amplitude potential
    tends to the linear one with a relative difference of second order."""
    rng = np.random.default_rng(3)
    base = O.ar_generate(rng.standard_normal((6, 16, 16)),
                         np.array([[[0, 0.5], [0.3, -0.2]], [[0.2, 0.1], [0.1, 0.05]]]))
    delta = (2 * np.pi,) * 3
    args = (12.0, (0.0, -10.0), (5.0, 0.0), (6, 3))
    diffs = []
    for eps in (1e-1, 1e-2, 1e-3):
        ha = O.high_amplitude_potential(eps * base, delta, *args)
        lin = O.linear_potential(eps * base, delta, *args)
        diffs.append(np.max(np.abs(ha - lin)) / np.max(np.abs(lin)))
    assert diffs[2] < 1e-5
    for k in range(2):
        assert 50 < diffs[k] / diffs[k + 1] < 200, diffs


def ha_potential_numpy(zeta, delta, depth, lb, ub, npts, conj=False):
    """numpy restatement of the high-amplitude solver with complex FFTs
    (conj=True plants a sign error in the imaginary part of D)."""
    nt, nx, ny = zeta.shape
    dtd = (ub[0] - lb[0]) / max(npts[0] - 1, 1)
    dzd = (ub[1] - lb[1]) / max(npts[1] - 1, 1)
    out = np.empty((npts[0], npts[1], nx, ny))
    for i in range(npts[0]):
        it = int(lb[0] + dtd * i)
        zt = O.derivative_t(zeta, delta[0], it)
        zx = np.gradient(zeta[it], delta[1], axis=0, edge_order=2)
        zy = np.gradient(zeta[it], delta[2], axis=1, edge_order=2)
        s = np.sqrt(1 + zx ** 2 + zy ** 2)
        den = 1j * ((zx + zy) / s - zx - zy) - 1 / s
        F = zt / (np.conj(den) if conj else den)
        kx = factor_waves_ref(zeta[it], int(delta[0] * it), delta[0])[0]
        ky = factor_waves_ref(zeta[it].T, int(delta[0] * it), delta[0])[0]
        G = np.fft.fft2(F)
        for j in range(npts[1]):
            w = O.window(nx, ny, 1 / kx, 1 / ky, lb[1] + dzd * j, depth)
            out[i, j] = edges_numpy(np.real(np.fft.ifft2(G * w)))
    return out


def test_high_amplitude_against_numpy():
    rng = np.random.default_rng(11)
    zeta = O.ar_generate(rng.standard_normal((6, 16, 32)),
                         np.array([[[0, 0.5], [0.3, -0.2]], [[0.2, 0.1], [0.1, 0.05]]]))
    delta = (2 * np.pi, 0.5, 0.5)
    args = (12.0, (0.0, -10.0), (5.0, 0.0), (6, 4))
    got = O.high_amplitude_potential(zeta, delta, *args)
    ref = ha_potential_numpy(zeta, delta, *args)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * scale)
    # the check sees the sign of Im D: it reaches the output only through the
    # odd part of the rotated window (about 1e-8 of the field here), far
    # above the agreement of the two restatements
    bad = ha_potential_numpy(zeta, delta, *args, conj=True)
    assert np.max(np.abs(bad - got)) > 1e-10 * scale
