"""Pins of the oracle's random-number layer against independent sources.

* MT19937 (src/libcilia/MersenneTwister.h): numpy's own MT19937 with the
  same init_genrand seeding, and the C++ standard's known answer (the
  10000th output of a default-seeded std::mt19937 is 4123659995).
* rand / randDblExc scalings (MersenneTwister.h:146-147, 158-159) against
  numpy's raw 32-bit stream.
* Philox4x64-10: numpy's Philox (which advances the counter before use) and
  the Random123 known answer for counter 0, key 0.
* Marsaglia-Tsang Gamma (mpcd.cpp:13-43): Kolmogorov-Smirnov against
  scipy's Gamma(3(N-1)/2) and its first two moments.
* Rotation axis (mpcd.cpp:276-284): uniform on the sphere.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O


def np_mt(seed):
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    return bg


def test_mt19937_known_answer_and_numpy():
    m = O.MTRand(5489)
    r = m.u32(10000)
    assert r[-1] == 4123659995
    ref = np_mt(5489).random_raw(10000)
    assert np.array_equal(r, ref.astype(np.uint64))
    m2 = O.MTRand(12345)
    assert np.array_equal(m2.u32(700), np_mt(12345).random_raw(700).astype(np.uint64))


def test_mt_scalings_against_numpy_raw():
    raw = np_mt(777).random_raw(4)
    m = O.MTRand(777)
    assert m.rand() == raw[0] * (1.0 / 4294967295.0)
    assert m.dblexc() == (raw[1] + 0.5) * (1.0 / 4294967296.0)
    x1 = (raw[2] + 0.5) / 4294967296.0
    x2 = (raw[3] + 0.5) / 4294967296.0
    assert abs(m.normal() - np.sqrt(-2 * np.log(x1)) * np.cos(2 * np.pi * x2)) < 1e-15


def test_philox_against_numpy_and_random123():
    assert [int(x) for x in O.philox4x64([0, 0, 0, 0], [0, 0])] == [
        0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]
    rng = np.random.default_rng(3)
    for _ in range(20):
        ctr = [int(x) for x in rng.integers(0, 2 ** 63, 4)]
        key = [int(x) for x in rng.integers(0, 2 ** 63, 2)]
        ref = np.random.Philox(counter=ctr, key=key).random_raw(4)
        c1 = list(ctr)
        c1[0] += 1  # numpy increments the counter before producing a block
        assert np.array_equal(O.philox4x64(c1, key), ref)


def test_philox4x32_known_answers():
    # Random123's kat_vectors for philox4x32_10 (ROCm's rocRAND engine gives
    # the same words)
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, out in kat:
        assert [int(x) for x in O.philox4x32(ctr, key)] == out


@pytest.mark.parametrize("nc", [2, 3, 10, 25])
def test_gamma_mt_distribution(nc):
    m = O.MTRand(100 + nc)
    k = 1.5 * (nc - 1)
    x = np.array([m.gamma(nc) for _ in range(20000)])
    assert stats.kstest(x, stats.gamma(k).cdf).pvalue > 1e-3
    assert abs(x.mean() - k) < 5 * np.sqrt(k / len(x))


@pytest.mark.parametrize("nc", [2, 10])
def test_gamma_philox_distribution(nc):
    k = 1.5 * (nc - 1)
    x = np.array([O.philox_gamma(9, 3, c, nc) for c in range(20000)])
    assert stats.kstest(x, stats.gamma(k).cdf).pvalue > 1e-3


def test_axis_uniform_on_sphere():
    m = O.MTRand(42)
    a = np.array([m.axis() for _ in range(20000)])
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0, atol=1e-12)
    assert stats.kstest(a[:, 2], stats.uniform(-1, 2).cdf).pvalue > 1e-3
    assert np.all(np.abs(a.mean(axis=0)) < 0.03)
    assert abs((a[:, 0] ** 2).mean() - 1 / 3) < 0.01
    p = np.array([O.philox_axis(1, 2, c) for c in range(20000)])
    assert stats.kstest(p[:, 2], stats.uniform(-1, 2).cdf).pvalue > 1e-3
    assert abs((p[:, 1] ** 2).mean() - 1 / 3) < 0.01


def test_philox_shift_grid():
    for step in range(50):
        a = O.philox_shift(7, step)
        assert np.all((a > 0) & (a < 1))
        q = a * 2 ** 24
        assert np.array_equal(q, np.round(q))
        # differences of shifts are exact in FP32
        b = O.philox_shift(7, step + 1)
        assert np.array_equal((b - a).astype(np.float32).astype(np.float64), b - a)


def test_fluid_create_philox_stream_and_distributions():
    """The oracle's Philox Fluid::create: its words are the Philox blocks of
    counters {i, 0, 0, 'FLUD'} / {i, 1, 0, 'FLUD'} (checked against the
    pinned Philox for a few particles by hand-made Box-Muller), positions
    uniform, velocities N(0, kT), mean removed."""
    seed, ext, n, kT = 11, (7.0, 5.0, 3.0), 60000, 1.5
    pos, vel = O.fluid_create_philox(seed, ext, n, kT)
    tag, key = 0x464C5544, [seed, 0x4D504344]
    raw = np.zeros((4, 3))
    for i in range(4):
        p = O.philox4x64([i, 0, 0, tag], key)
        q = O.philox4x64([i, 1, 0, tag], key)
        u = lambda x: ((int(x) >> 11) + 0.5) / 2.0 ** 53
        assert np.array_equal(pos[i], [u(p[k]) * ext[k] for k in range(3)])
        r0, r1 = np.sqrt(-2 * np.log(u(q[0]))), np.sqrt(-2 * np.log(u(q[2])))
        raw[i] = np.array([r0 * np.cos(2 * np.pi * u(q[1])), r0 * np.sin(2 * np.pi * u(q[1])),
                           r1 * np.cos(2 * np.pi * u(q[3]))]) * np.sqrt(kT)
    shift = raw - vel[:4, :3]  # the removed mean, the same for every particle
    assert np.abs(shift - shift[0]).max() < 1e-14 and np.abs(shift[0]).max() < 0.02
    for a in range(3):
        assert stats.kstest(pos[:, a] / ext[a], "uniform").pvalue > 1e-3
        assert stats.kstest(vel[:, a] / np.sqrt(kT), "norm").pvalue > 1e-3
    assert np.abs(vel[:, :3].mean(axis=0)).max() < 1e-14
    assert np.all(vel[:, 3] == 1.0)
