"""The OpenMP bucket step that bench.py times as the CPU baseline
(orc_omp_*, the reference's threaded algorithm: mpcd.cpp:140-267,
fluid.cpp:152-167) is the oracle's step: with one thread it draws from one
MT19937 in the reference's order (shift, then per cell Gamma and axis, then
the next shift) and its solvent after two steps is bit-identical to
orc_collide + orc_stream fed by the same generator.  With several threads
each thread draws from its own generator, so only the invariants hold:
momentum to roundoff and the thermostat's temperature."""
import numpy as np

import oracle as O


def _fluid(L, seed):
    rng = np.random.default_rng(seed)
    n = 10 * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L, dtype=np.float64)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    return pos, vel


def test_one_thread_equals_oracle_collide():
    L, seed, h = (10, 9, 8), 77, 0.1
    pos, vel = _fluid(L, 1)
    p1, v1 = pos.copy(), vel.copy()
    sim = O.OmpSim(L, p1, v1, threads=1, seed=seed)
    sim.steps(2, h)
    sim.close()
    mt = O.MTRand(seed)
    a = np.array([mt.dblexc() for _ in range(3)])
    for _ in range(2):
        cfg, keep, _ = O.make_cfg(L, rng=O.RNG_MT, mt=mt)
        O.collide(cfg, a, pos, vel)
        a = np.array([mt.dblexc() for _ in range(3)])
        O.stream(pos, vel, h)
    assert np.array_equal(p1, pos) and np.array_equal(v1, vel)


def test_threads_and_fast_math_build_keep_invariants():
    L, h = (16, 16, 16), 0.1
    pos, vel = _fluid(L, 2)
    p0 = vel[:, :3].sum(0)
    for baseline in (False, True):
        p, v = pos.copy(), vel.copy()
        sim = O.OmpSim(L, p, v, threads=4, seed=3, baseline=baseline)
        sim.steps(10, h)
        sim.close()
        assert np.abs(v[:, :3].sum(0) - p0).max() < 1e-9
        assert abs(np.mean(v[:, :3] ** 2) - 1.0) < 0.02
        assert np.array_equal(p, pos + 0) is False
