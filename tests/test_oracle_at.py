"""Pins of the oracle's MPC-AT collision rules (absent in the reference:
Noguchi, Kikuchi, Gompper, EPL 78 (2007) 10005; Gompper et al., Adv.
Polym. Sci. 221 (2009) 1, eq. 20).  Parity unpinned by the reference, so
the rule is pinned by what the physics fixes:

* every cell's linear momentum is conserved, with mixed masses (beads);
* with +a every cell's angular momentum about its centre of mass (in the
  cell-local coordinates the rule uses) is conserved, and without +a it is
  not;
* the relative velocities are thermal: E[sum m |v' - u|^2] = 3 (N - 1) kT
  per cell of N particles after one AT collision from any state;
* omega solves I omega = L (numpy's least-squares solution, also for the
  collinear two-particle cell);
* injected normals reproduce the Philox and MT19937 runs (draw order: per
  cell ascending, per particle in insertion order, three normals).
"""
import numpy as np
import pytest

import oracle as O

H = 0.1


def _state(L, rng, rho=10, nb=0):
    n = rho * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    bpos = rng.uniform(0, 1, (nb, 3)) * np.array(L)
    bvel = np.full((nb, 4), 5.0)
    bvel[:, :3] = rng.normal(size=(nb, 3)) * 0.5
    return pos, vel, bpos, bvel


def _local(x, a):
    t = x + a
    return (t - np.floor(t)).astype(np.float32).astype(np.float64)


def _per_cell(L, a, pos, vel, bpos, bvel):
    """Per cell: mass, momentum, angular momentum about the centre of mass
    of the cell-local coordinates."""
    cells = np.concatenate([O.box(L, a, pos), O.box(L, a, bpos) if len(bpos) else np.zeros(0, np.int32)])
    x = np.vstack([pos, bpos])
    v = np.vstack([vel, bvel])
    m = v[:, 3]
    s = _local(x, a)
    nc = int(np.prod(L))
    M = np.bincount(cells, m, nc)
    P = np.stack([np.bincount(cells, m * v[:, k], nc) for k in range(3)], 1)
    C = np.stack([np.bincount(cells, m * s[:, k], nc) for k in range(3)], 1) / np.maximum(M, 1e-300)[:, None]
    r = s - C[cells]
    Lm = np.stack([np.bincount(cells, x_, nc) for x_ in (m[:, None] * np.cross(r, v[:, :3])).T], 1)
    return M, P, Lm


@pytest.mark.parametrize("rule", [O.AT, O.AT_A])
def test_cell_momentum_conserved(rule):
    L = (6, 5, 4)
    rng = np.random.default_rng(1)
    pos, vel, bpos, bvel = _state(L, rng, nb=60)
    a = rng.uniform(0, 1, 3)
    M0, P0, L0 = _per_cell(L, a, pos, vel, bpos, bvel)
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=5, step=3, rule=rule)
    O.collide(cfg, a, pos, vel, bpos, bvel)
    M1, P1, L1 = _per_cell(L, a, pos, vel, bpos, bvel)
    np.testing.assert_allclose(P1, P0, rtol=0, atol=1e-11)
    if rule == O.AT_A:
        np.testing.assert_allclose(L1, L0, rtol=0, atol=1e-11)
    else:
        assert np.abs(L1 - L0).max() > 0.1


def test_relative_velocities_thermal():
    """AT without +a: E[sum_i m |v'_i - u|^2] = 3 (N - 1) kT per cell."""
    L = (16, 16, 16)
    rng = np.random.default_rng(2)
    pos, vel, _, _ = _state(L, rng)
    vel[:, :3] *= 3.0  # far from kT = 1 before the collision
    a = rng.uniform(0, 1, 3)
    kT = 1.7
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=9, step=0, rule=O.AT, kT=kT)
    O.collide(cfg, a, pos, vel)
    cells = O.box(L, a, pos)
    nc = int(np.prod(L))
    N = np.bincount(cells, minlength=nc)
    u = np.stack([np.bincount(cells, vel[:, k], nc) for k in range(3)], 1) / np.maximum(N, 1)[:, None]
    e = np.bincount(cells, ((vel[:, :3] - u[cells]) ** 2).sum(1), nc)
    ok = N > 1
    ratio = e[ok].sum() / (3.0 * (N[ok] - 1).sum() * kT)
    assert abs(ratio - 1.0) < 0.01  # about 120k degrees of freedom: sigma ~ 0.004


def test_omega_solves_inertia_system():
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = rng.integers(3, 12)
        r = rng.uniform(-0.5, 0.5, (n, 3))
        m = rng.choice([1.0, 5.0], n)
        r -= (m[:, None] * r).sum(0) / m.sum()
        I = sum(mi * (ri @ ri * np.eye(3) - np.outer(ri, ri)) for mi, ri in zip(m, r))
        Lv = rng.normal(size=3)
        w = O.at_omega([I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]], Lv)
        np.testing.assert_allclose(I @ w, Lv, rtol=1e-9, atol=1e-9)
    # two particles: I is singular; L (normal to the pair) is matched by the
    # minimum-norm solution
    r = np.array([[0.1, -0.2, 0.3], [-0.1, 0.2, -0.3]])
    I = sum(ri @ ri * np.eye(3) - np.outer(ri, ri) for ri in r)
    Lv = np.cross(r[0], rng.normal(size=3))
    w = O.at_omega([I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]], Lv)
    np.testing.assert_allclose(w, np.linalg.lstsq(I, Lv, rcond=None)[0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("rule", [O.AT, O.AT_A])
def test_injection_reproduces_philox_and_mt(rule):
    L = (5, 4, 3)
    rng = np.random.default_rng(4)
    pos, vel, bpos, bvel = _state(L, rng, nb=20)
    a = rng.uniform(0, 1, 3)
    n = len(pos) + len(bpos)
    # Philox run, recorded normals, then the same normals injected
    v1, b1 = vel.copy(), bvel.copy()
    cfg, keep, rec = O.make_cfg(L, rng=O.RNG_PHILOX, seed=11, step=2, rule=rule, record=True, n_record_ran=n)
    O.collide(cfg, a, pos, v1, bpos, b1)
    cells = np.concatenate([O.box(L, a, pos), O.box(L, a, bpos)])
    _, _, rank = O.bin_cells(cells, int(np.prod(L)))
    for i in rng.choice(n, 20, replace=False):
        assert rec[3][3 * i + 1] == O.philox_ran(11, 2, int(cells[i]), int(rank[i]), 1)
    v2, b2 = vel.copy(), bvel.copy()
    cfg, keep2, _ = O.make_cfg(L, rng=O.RNG_INJECT, rule=rule, ran=rec[3])
    O.collide(cfg, a, pos, v2, bpos, b2)
    assert np.array_equal(v1, v2) and np.array_equal(b1, b2)
    # MT19937: three normals per particle, cells ascending, insertion order
    mt = O.MTRand(77)
    cfg, keep3, rec3 = O.make_cfg(L, rng=O.RNG_MT, mt=mt, rule=rule, record=True, n_record_ran=n)
    O.collide(cfg, a, pos, vel.copy(), bpos, bvel.copy())
    replay = O.MTRand(77)
    order = np.lexsort((np.arange(n), cells))
    for i in order:
        for k in range(3):
            assert rec3[3][3 * i + k] == replay.normal()
