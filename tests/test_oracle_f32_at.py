"""The oracle's MPC-AT(+a) in the FP32-bulk form (orc_f32_step with rule
AT / AT_A), pinned:

* against the FP64 form's MPC-AT (orc_collide, itself pinned in
  test_oracle_at.py) with the same injected normals, on the same positions:
  the velocities differ only by the FP32 grid (random velocities on 2^-21,
  lever arms on 2^-24, FP32 update, the residual): within 1e-5;
* fluid momentum exactly conserved (an integer identity);
* MPC-AT+a conserves each cell's angular momentum about its centre of mass
  to FP32 rounding, and plain MPC-AT does not (a dropped or mis-signed
  omega x r fails);
* the result does not depend on the order of the particles, with Philox
  normals named by each particle's offsets (the exact-integer sums).
"""
import numpy as np
import pytest

import oracle as O

H = 0.1


def _fluid(L, rng, n_per=10):
    n = n_per * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    return pos, vel


def _collide_f32(L, a, pos, vel, rule, ran=None, bpos=None, bvel=None, seed=0, ids=True):
    cell, off, w = O.f32_encode(L, a, pos, vel)
    before = (cell.copy(), off.copy(), w.copy())
    if ran is None:
        cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=3, rule=rule)
    else:
        cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, rule=rule, ran=ran)
    bv = None if bvel is None else bvel.copy()
    O.f32_step(cfg, a, a, 0.0, cell, off, w, bpos3=bpos, bvel4=bv, ids=np.arange(len(cell), dtype=np.int32) if ids else None)
    return before, (cell, off, w), bv


@pytest.mark.parametrize("rule_name", ["AT", "AT_A"])
def test_f32_at_matches_fp64_at(rule_name):
    L = (6, 5, 4)
    rng = np.random.default_rng(11)
    pos, vel = _fluid(L, rng)
    nb = 40
    bpos = rng.uniform(0, 1, (nb, 3)) * np.array(L)
    bvel = np.full((nb, 4), 5.0)
    bvel[:, :3] = rng.normal(size=(nb, 3)) * 0.4
    a = rng.uniform(0, 1, 3)
    # both forms see the same positions: the FP32 records decoded
    cell, off, w = O.f32_encode(L, a, pos, vel)
    p32, v32 = O.f32_decode(L, a, cell, off, w)
    n = len(pos)
    ran = rng.normal(size=3 * (n + nb))
    rule = getattr(O, rule_name)
    _, (c2, o2, w2), bv32 = _collide_f32(L, a, p32, v32, rule, ran=ran, bpos=bpos, bvel=bvel)
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, rule=rule, ran=ran)
    v64 = v32.copy()
    bv64 = bvel.copy()
    O.collide(cfg, a, p32, v64, bpos, bv64)
    assert np.abs(w2 - v64[:, :3]).max() < 1e-5
    assert np.abs(bv32[:, :3] - bv64[:, :3]).max() < 1e-5


@pytest.mark.parametrize("rule_name", ["AT", "AT_A"])
def test_f32_at_momentum_exact(rule_name):
    L = (5, 5, 5)
    rng = np.random.default_rng(12)
    pos, vel = _fluid(L, rng)
    a = rng.uniform(0, 1, 3)
    (c0, o0, w0), (c1, o1, w1), _ = _collide_f32(L, a, pos, vel, getattr(O, rule_name), seed=4)
    q0 = np.rint(w0.astype(np.float64) * 2 ** 21).astype(np.int64)
    q1 = np.rint(w1.astype(np.float64) * 2 ** 21).astype(np.int64)
    assert np.array_equal(q0.sum(0), q1.sum(0))
    assert np.abs(w1 - w0).max() > 0.1  # the collision did something


def _cell_angular_momentum(cell, off, w):
    """Per cell: sum (r - C) x v with r the lever arm on the 2^-24 grid."""
    r = np.rint(off.astype(np.float64) * 2 ** 24) / 2 ** 24
    v = w.astype(np.float64)
    out = {}
    for c in np.unique(cell):
        m = cell == c
        rc = r[m] - r[m].mean(0)
        out[c] = np.cross(rc, v[m]).sum(0)
    return out


def test_f32_at_a_conserves_cell_angular_momentum():
    L = (4, 4, 4)
    rng = np.random.default_rng(13)
    pos, vel = _fluid(L, rng)
    a = rng.uniform(0, 1, 3)
    res = {}
    for name in ("AT", "AT_A"):
        (c0, o0, w0), (c1, o1, w1), _ = _collide_f32(L, a, pos, vel, getattr(O, name), seed=9)
        # the collision keeps the cells: compare per cell before and after it (stream of 0)
        assert np.array_equal(c0, c1) and np.array_equal(o0, o1)
        L0 = _cell_angular_momentum(c0, o0, w0)
        L1 = _cell_angular_momentum(c1, o1, w1)
        res[name] = max(np.abs(L1[c] - L0[c]).max() for c in L0)
    assert res["AT_A"] < 2e-5  # FP32 update and the residual's velocity grid step
    assert res["AT"] > 0.1


def test_f32_at_order_independent():
    L = (4, 5, 3)
    rng = np.random.default_rng(14)
    pos, vel = _fluid(L, rng)
    a = rng.uniform(0, 1, 3)
    _, (c1, o1, w1), _ = _collide_f32(L, a, pos, vel, O.AT_A, seed=21, ids=False)
    perm = rng.permutation(len(pos))
    _, (c2, o2, w2), _ = _collide_f32(L, a, pos[perm], vel[perm], O.AT_A, seed=21, ids=False)
    inv = np.argsort(perm)
    assert np.array_equal(c2[inv], c1) and np.array_equal(o2[inv], o1)
    # every particle but the residual owners (one per cell, chosen by content) equal
    assert np.array_equal(w2[inv], w1)


def test_f32_at_temperature():
    """E[sum over the cell |v' - u'|^2] = 3 (N - 1) kT after one collision
    from any start (the Andersen rule draws the relative velocities)."""
    L = (8, 8, 8)
    rng = np.random.default_rng(15)
    pos, vel = _fluid(L, rng)
    vel[:, :3] *= 0.3  # cold start
    a = rng.uniform(0, 1, 3)
    _, (c1, o1, w1), _ = _collide_f32(L, a, pos, vel, O.AT, seed=5, ids=False)
    tot, dof = 0.0, 0
    for c in np.unique(c1):
        m = c1 == c
        v = w1[m].astype(np.float64)
        tot += ((v - v.mean(0)) ** 2).sum()
        dof += 3 * (m.sum() - 1)
    assert abs(tot / dof - 1.0) < 0.02
