"""Pins the oracle's Mersenne Twister and normal transform.

The dynamic-creator generator (genrand_mt / sgenrand_mt, called at
src/parallel_mt.hh:236,259) run with the MT19937 parameter set must equal
std::mt19937; the libstdc++ polar normal (random.tcc:1802-1836) over it must
equal std::normal_distribution bit for bit.
"""
import os

import numpy as np
import pytest

import pyoracle as O


@pytest.fixture(scope="module")
def gold(golden_dir):
    return np.load(os.path.join(golden_dir, "libstdcxx_random.npz"))


def test_mt19937_known_answer():
    # The C++ standard fixes the 10000th draw of a default-seeded mt19937.
    draws = O.mt_fill(O.mt19937_params(), 5489, 10000)
    assert int(draws[-1]) == 4123659995


def test_mt19937_golden(gold):
    for s in gold["seeds"]:
        ref = gold[f"mt19937_{int(s)}"]
        got = O.mt_fill(O.mt19937_params(), int(s), ref.size)
        np.testing.assert_array_equal(got, ref)


def test_golden_default_10000th(gold):
    assert int(gold["mt19937_default_10000th"][0]) == 4123659995


@pytest.mark.parametrize("si", [0, 1])
def test_normal_double_bit_exact(gold, si):
    sigma = float(gold["sigmas"][si])
    for s in gold["seeds"]:
        ref = gold[f"normal_d_{int(s)}_{si}"]
        got = O.noise_sequential(O.mt19937_params(), int(s), sigma, ref.size)
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("si", [0, 1])
def test_normal_float_bit_exact(gold, si):
    sigma = float(gold["sigmas"][si])
    for s in gold["seeds"]:
        ref = gold[f"normal_f_{int(s)}_{si}"]
        got = O.noise_sequential(O.mt19937_params(), int(s), np.float32(sigma), ref.size,
                                 dtype=np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_normal_moments():
    x = O.noise_sequential(O.mt19937_params(), 7, 2.0, 200000)
    assert abs(x.mean()) < 0.02
    assert abs(x.var() - 4.0) < 0.06


def test_partitioned_noise_layout():
    """Stream p fills partition p row-major; a fresh normal per partition."""
    shape, part = (6, 5, 4), (4, 3, 4)
    nparts = 2 * 2 * 1
    plist = [O.mt19937_params() for _ in range(nparts)]
    seeds = np.array([3, 9, 27, 81], np.uint32)
    z = O.noise_partitioned(plist, seeds, shape, part, 1.5)
    pid = 0
    for a in range(2):
        for b in range(2):
            lo = (a * 4, b * 3, 0)
            hi = (min(lo[0] + 4, 6), min(lo[1] + 3, 5), 4)
            n = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])
            ref = O.noise_sequential(O.mt19937_params(), int(seeds[pid]), 1.5, n)
            got = z[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].reshape(-1)
            np.testing.assert_array_equal(got, ref)
            pid += 1


def test_partitioned_noise_too_few_streams():
    with pytest.raises(RuntimeError):
        O.noise_partitioned([O.mt19937_params()], None, (4, 4, 4), (2, 2, 2), 1.0)
