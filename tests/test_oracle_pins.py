"""Known answers for the two oracle functions that no closed form of the
whole solver reaches: the edge fix-up and the wave-number scan.

Edge fix-up (src/velocity/basic_solver.cc:61-68, plane of
src/interpolate.cc:3-29): each fixed point is the plane through three grid
points evaluated at a fourth.  On integer data the plane values are exact, so
the expected arrays below are worked out by hand:

  res(0,0) = 3 f(1,1) - f(2,1) - f(1,2)
  res(k,0) = f(k,1) + f(k-1,1) - f(k-1,2)            k = 1 .. nx-1
  res(0,l) = f(1,l) + f(1,l-1) - f(2,l-1)            l = 1 .. ny-1

in that order, so res(0,1) reads the already fixed res(1,0) and res(2,0).

Wave-number scan (src/factor_waves.cc:22-74): an event is a position x in
1 .. n-2 whose differences dw1 = z(x)-z(x-1), dw2 = z(x)-z(x+1) are both
positive or both non-positive; its vertex is that of the parabola through
(x-1, x, x+1) placed at slice t: period (t + s) dt with
s = (z2 - z1) / (2 (2 z0 - z1 - z2)), elevation z0 + (z2 - z1)^2 /
(8 (2 z0 - z1 - z2)).  Events form runs of equal elevation sign (scanned
column by column); a run is represented by its first event of largest
|elevation|; when run r+1 starts, the pair (r, r-1) is scored as
k = (2 pi / |2 (T_r - T_{r-1})|)^2 / 9.82, so the last run is never scored.
The slices below are built so that each rule decides the answer.
"""
import math

import numpy as np
import pytest

import pyoracle as O

G = 9.82
DBL_MAX = np.finfo(np.float64).max
DBL_MIN = np.finfo(np.float64).tiny

# ---- edge fix-up ----

FIELD = np.array([[5, -2, 12, 1, 4],
                  [3, 8, -1, 6, 2],
                  [-4, 0, 9, -3, 5],
                  [6, 2, -5, 4, 1]], dtype=np.float64)

# worked by hand from the formulas in the module docstring:
#   res(0,0) = 3*8 - 0 - (-1)      = 25
#   res(1,0) = 8 + (-2) - 12       = -6
#   res(2,0) = 0 + 8 - (-1)        = 9
#   res(3,0) = 2 + 0 - 9           = -7
#   res(0,1) = 8 + (-6) - 9        = -7    (reads the fixed res(1,0), res(2,0))
#   res(0,2) = -1 + 8 - 0          = 7
#   res(0,3) = 6 + (-1) - 9        = -4
#   res(0,4) = 2 + 6 - (-3)        = 11
EXPECTED = FIELD.copy()
EXPECTED[0, :] = [25, -7, 7, -4, 11]
EXPECTED[1:, 0] = [-6, 9, -7]


def test_edges_known_answer():
    got = O.interp_edges(FIELD)
    np.testing.assert_array_equal(got, EXPECTED)


def _edges_variant(f, order="cols_first", col_pts=((-1, 1), (0, 1), (-1, 2))):
    """Plausible misreadings of the fix-up, for the check that the known
    answer tells them apart."""
    r = f.copy()
    nx, ny = r.shape

    def plane(p1, p2, p3, x):
        # solve for f = c0 x0 + c1 x1 + c2 through three points
        A = np.array([[p[0], p[1], 1.0] for p in (p1, p2, p3)])
        c = np.linalg.solve(A, [r[p] for p in (p1, p2, p3)])
        return c[0] * x[0] + c[1] * x[1] + c[2]

    def cols():
        for k in range(1, nx):
            pts = [(k + a, b) for a, b in col_pts]
            r[k, 0] = plane(*pts, (k, 0))

    def rows():
        for l in range(1, ny):
            r[0, l] = plane((1, l - 1), (1, l), (2, l - 1), (0, l))

    r[0, 0] = plane((1, 1), (1, 2), (2, 1), (0, 0))
    if order == "cols_first":
        cols()
        rows()
    else:
        rows()
        cols()
    return r


def test_edges_variants_are_distinguished():
    # the linear-algebra restatement in the reference's order gives the
    # hand-worked answer ...
    np.testing.assert_allclose(_edges_variant(FIELD), EXPECTED, atol=1e-12)
    # ... and each misreading does not
    assert np.max(np.abs(_edges_variant(FIELD, order="rows_first") - EXPECTED)) > 1
    swapped = _edges_variant(FIELD, col_pts=((-1, 1), (0, 1), (0, 2)))
    assert np.max(np.abs(swapped - EXPECTED)) > 1


def test_edges_plane_is_fixed_point():
    # on a plane every fix-up reproduces the plane exactly
    x = np.arange(7)[:, None]
    y = np.arange(9)[None, :]
    f = 3.0 * x - 2.0 * y + 5.0 + 0.0 * (x + y)
    np.testing.assert_array_equal(O.interp_edges(f), f)


# ---- wave-number scan ----

# One line of 20 points.  Events (x: triple z(x-1), z(x), z(x+1) -> s, elevation):
#   x=1:  ( 0,   4,   2 )   s =  1/6,   elev  49/12   run 0 (+)
#   x=4:  (-1,  -2,  -1.5)  s =  1/6,   elev -97/48   run 1 (-)
#   x=6:  (-1.5,-1.2,-4 )   s = -25/62, elev -0.948   run 1, smaller
#   x=7:  (-1.2,-4,  -3 )   s =  9/38,  elev -4.107   run 1, larger: represents it
#   x=10: ( 4,   5,   1 )   s = -3/10,  elev  5.225   run 2 (+)
#   x=12: ( 1,   0.5, 1 )   s =  0,     elev  0.5     run 2, smaller
#   x=14: ( 1,   5,   4 )   s =  3/10,  elev  5.225   run 2, a tie: x=10 stays
#   x=17: ( 0.5,-2,  -1.5)  s =  1/3,   elev -13/6    run 3 (-), never scored
# Scored: (run 1, run 0): dT = (9/38 - 1/6) dt = (4/57) dt
#         (run 2, run 1): dT = (-3/10 - 9/38) dt = -(51/95) dt
LINE = np.array([0, 4, 2, -1, -2, -1.5, -1.2, -4, -3, 4, 5, 1, 0.5, 1, 5, 4, 0.5, -2,
                 -1.5, 0], dtype=np.float64)


def k_of(ds, dt):
    w = 2 * math.pi / abs(2 * ds * dt)
    return w * w / G


@pytest.mark.parametrize("t,dt", [(0, 1.0), (3, 2 * math.pi), (250, 2 * math.pi)])
def test_factor_waves_known_answer(t, dt):
    kmin = k_of(51 / 95, dt)
    kmax = k_of(4 / 57, dt)
    # along x (one column), and the same line along y (one row)
    got = O.factor_waves(LINE[:, None], t, dt)
    np.testing.assert_allclose(got[0], kmin, rtol=1e-12)
    np.testing.assert_allclose(got[2], kmax, rtol=1e-12)
    # one column: no position inside the line along y, so no events there
    assert got[1] == DBL_MAX and got[3] == DBL_MIN
    got = O.factor_waves(LINE[None, :], t, dt)
    np.testing.assert_allclose(got[1], kmin, rtol=1e-12)
    np.testing.assert_allclose(got[3], kmax, rtol=1e-12)
    assert got[0] == DBL_MAX and got[2] == DBL_MIN


def test_factor_waves_rules_decide():
    """The known answer above rejects each plausible misreading."""
    dt = 1.0
    kmin, kmax = k_of(51 / 95, dt), k_of(4 / 57, dt)
    # representative = last event of largest |elevation|: run 2 would be x=14
    assert not math.isclose(k_of(3 / 10 - 9 / 38, dt), kmin, rel_tol=1e-6)
    assert not math.isclose(k_of(3 / 10 - 9 / 38, dt), kmax, rel_tol=1e-6)
    # no replacement by larger events: run 1 would be x=4, whose period
    # equals run 0's (k infinite), and (run 2, run 1) would differ by 7/15
    assert not math.isclose(k_of(7 / 15, dt), kmin, rel_tol=1e-6)
    # scoring the last run too would add (run 3, run 2): dT = 1/3 + 3/10
    assert k_of(1 / 3 + 3 / 10, dt) < kmin


def test_factor_waves_runs_cross_columns():
    """Two columns holding the same line: the scan runs column after column,
    so column 1's first run follows column 0's last run.  Runs:
    0:+1/6, 1:9/38, 2:-3/10, 3:1/3, 4:+1/6 (column 1), 5:9/38, 6:-3/10,
    7:1/3 (last, unscored).  Scored differences: 4/57, 51/95, 19/30 (3,2),
    1/6 (4,3), 4/57, 51/95."""
    dt = 2 * math.pi
    got = O.factor_waves(np.stack([LINE, LINE], axis=1), 5, dt)
    np.testing.assert_allclose(got[0], k_of(19 / 30, dt), rtol=1e-12)
    np.testing.assert_allclose(got[2], k_of(4 / 57, dt), rtol=1e-12)


def test_factor_waves_no_events():
    """A monotone slice has no events at all (the reference reads events[0]
    of an empty vector there; the oracle and the GPU keep the initial
    bounds: the largest double and the smallest normal double)."""
    z = np.add.outer(np.arange(6.0), 2 * np.arange(5.0))
    assert O.factor_waves(z, 4, 2 * math.pi) == (DBL_MAX, DBL_MAX, DBL_MIN, DBL_MIN)


def test_factor_waves_fewer_than_three_runs():
    """Two runs only: the first switch scores nothing and the last run is
    never scored, so the bounds stay at their initial values."""
    z = np.array([0.0, 3.0, 1.0, -2.0, 0.0])  # events x=1 (+), x=3 (-)
    got = O.factor_waves(z[:, None], 0, 1.0)
    assert got[0] == DBL_MAX and got[2] == DBL_MIN
