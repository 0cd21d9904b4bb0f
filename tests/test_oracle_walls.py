"""Pins the oracle's planar walls (SLIP / NOSLIP) against what fixes them
independently of the oracle's own code:

* bounce-back (mpcd.cpp:346-377): geometry.  NOSLIP is a reflection of the
  velocity relative to the moving top wall at the crossing time, SLIP a
  mirror in the top plane, the bottom wall reverses the whole velocity at
  the crossing time; particles that cross no wall stream freely.
* virtual particles (mpcd.cpp:175-198): rho - num virtual particles of
  total momentum vi enter the cell's sums exactly as rho - num real unit
  masses of velocity vi / (rho - num) would (their momentum vi, their
  kinetic energy |vi|^2 / (rho - num), their count and mass), so a wall
  collision must give the real particles the velocities of a wall-less
  collision with those particles added.
* which cells get them: the layer test as written in the reference
  (cid < Lx Ly or cid > Lx Ly (Lz - 1), so the first top-layer cell does
  not), and only when 0 < num < rho.
* draw order with MT19937: per cell [3 normals] -> [gamma] -> axis.
"""
import numpy as np
import pytest

import oracle as O

TS = 0.1


def _free(pos, vel, ts):
    return pos + vel[:, :3] * ts


@pytest.mark.parametrize("vw", [(0.0, 0.0), (0.3, -0.2)])
def test_noslip_bounce_is_reflection_about_wall_velocity(vw):
    rng = np.random.default_rng(1)
    Lz = 6
    n = 4000
    pos = rng.uniform(0.0, Lz - 1.0, (n, 3))
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(0.0, 3.0, (n, 3))
    p0, v0 = pos.copy(), vel.copy()
    O.stream_walls(pos, vel, TS, Lz, 1, vw)
    free = _free(p0, v0, TS)
    top = free[:, 2] > Lz - 1
    bot = free[:, 2] < 0
    assert top.sum() > 50 and bot.sum() > 50
    none = ~top & ~bot
    np.testing.assert_array_equal(pos[none], free[none])
    np.testing.assert_array_equal(vel[none], v0[none])
    w = np.array([vw[0], vw[1], 0.0])
    # top: hit at t_hit, then v' = 2 w - v (z: -vz) for the remaining time
    t_hit = (Lz - 1 - p0[top, 2]) / v0[top, 2]
    v_new = 2 * w - v0[top, :3]
    exp = p0[top] + v0[top, :3] * t_hit[:, None] + v_new * (TS - t_hit)[:, None]
    np.testing.assert_allclose(pos[top], exp, rtol=0, atol=1e-12)
    np.testing.assert_allclose(vel[top, :3], v_new, rtol=0, atol=1e-15)
    # bottom: the whole velocity reversed at the crossing
    t_hit = -p0[bot, 2] / v0[bot, 2]
    exp = p0[bot] + v0[bot, :3] * t_hit[:, None] - v0[bot, :3] * (TS - t_hit)[:, None]
    np.testing.assert_allclose(pos[bot], exp, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(vel[bot, :3], -v0[bot, :3])
    assert pos[:, 2].min() >= 0 and pos[:, 2].max() <= Lz - 1


def test_slip_top_is_mirror_bottom_reverses():
    rng = np.random.default_rng(2)
    Lz = 5
    n = 4000
    pos = rng.uniform(0.0, Lz - 1.0, (n, 3))
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(0.0, 3.0, (n, 3))
    p0, v0 = pos.copy(), vel.copy()
    O.stream_walls(pos, vel, TS, Lz, 2, (0.7, 0.7))  # wall velocity plays no part in SLIP
    free = _free(p0, v0, TS)
    top = free[:, 2] > Lz - 1
    bot = free[:, 2] < 0
    assert top.sum() > 50 and bot.sum() > 50
    np.testing.assert_array_equal(pos[top, :2], free[top, :2])
    np.testing.assert_allclose(pos[top, 2], 2.0 * (Lz - 1) - free[top, 2], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(vel[top, :2], v0[top, :2])
    np.testing.assert_array_equal(vel[top, 2], -v0[top, 2])
    t_hit = -p0[bot, 2] / v0[bot, 2]
    exp = p0[bot] + v0[bot, :3] * t_hit[:, None] - v0[bot, :3] * (TS - t_hit)[:, None]
    np.testing.assert_allclose(pos[bot], exp, rtol=0, atol=1e-12)
    # NOSLIP and SLIP agree at the bottom wall
    pos2, vel2 = p0.copy(), v0.copy()
    O.stream_walls(pos2, vel2, TS, Lz, 1, (0.0, 0.0))
    np.testing.assert_array_equal(pos2[bot], pos[bot])


def _cell_center(L, a, cid):
    """A position inside cell cid under shift a (box(): floor(x + a))."""
    i = np.array([cid % L[0], (cid // L[0]) % L[1], cid // (L[0] * L[1])], dtype=np.float64)
    return i + 0.5 - np.asarray(a)


def _wall_layer(L, cid):
    layer = L[0] * L[1]
    return cid < layer or cid > layer * (L[2] - 1)


@pytest.mark.parametrize("thermostat", [1, 0])
def test_virtual_particles_equal_added_real_particles(thermostat):
    L = (4, 3, 5)
    rho = 10
    seed, step = 99, 4
    rng = np.random.default_rng(3)
    a = np.array([0.3, 0.6, 0.45])
    ncells = int(np.prod(L))
    # cells get 1..14 particles, placed inside their cells
    counts = rng.integers(1, 15, ncells)
    pos = np.vstack([_cell_center(L, a, c) + rng.uniform(-0.45, 0.45, (k, 3)) for c, k in enumerate(counts)])
    assert np.array_equal(np.bincount(O.box(L, a, pos), minlength=ncells), counts)
    vel = np.ones((len(pos), 4))
    vel[:, :3] = rng.normal(size=(len(pos), 3))
    # (a) walls, virtual particles from the Philox stream
    pa, va = pos.copy(), vel.copy()
    cfg, keep, rec = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=step, thermostat=thermostat, walls=1,
                                rho=rho, record=True)
    O.collide(cfg, a, pa, va)
    # (b) no walls; rho - num real particles of velocity vi / (rho - num)
    extra_p, extra_v = [], []
    for c in range(ncells):
        if _wall_layer(L, c) and counts[c] < rho:
            md = rho - counts[c]
            nr = np.array([O.philox_virt(seed, step, c, k) for k in range(3)])
            np.testing.assert_array_equal(rec[2][3 * c:3 * c + 3], nr)
            vi = np.sqrt(md * 1.0) * nr
            extra_p.append(np.repeat(_cell_center(L, a, c)[None], md, axis=0))
            v = np.ones((md, 4))
            v[:, :3] = vi / md
            extra_v.append(v)
        else:
            assert np.isnan(rec[2][3 * c])
    assert len(extra_p) > 10
    pb = np.vstack([pos] + extra_p)
    vb = np.vstack([vel] + extra_v)
    cfg_b, keep_b, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=step, thermostat=thermostat)
    O.collide(cfg_b, a, pb, vb)
    np.testing.assert_allclose(va[:, :3], vb[:len(pos), :3], rtol=0, atol=1e-11)


def test_virtual_particle_cells_follow_the_reference_layer_test():
    L = (2, 2, 3)
    a = np.array([0.5, 0.5, 0.5])
    ncells = int(np.prod(L))
    pos = np.vstack([_cell_center(L, a, c) for c in range(ncells)] + [_cell_center(L, a, 5)] * 20)
    vel = np.ones((len(pos), 4))
    vel[:, :3] = np.random.default_rng(4).normal(size=(len(pos), 3))
    cfg, keep, rec = O.make_cfg(L, rng=O.RNG_PHILOX, seed=1, step=0, walls=2, rho=10, record=True)
    O.collide(cfg, a, pos, vel)
    got = [c for c in range(ncells) if not np.isnan(rec[2][3 * c])]
    # bottom layer 0..3; top layer 8..11 except 8 (strict ">" in mpcd.cpp:177);
    # the middle layer 4..7 never
    assert got == [0, 1, 2, 3, 9, 10, 11]
    # a full wall cell gets none: fill cell 1 up to rho
    pos = np.vstack([pos] + [_cell_center(L, a, 1)] * 9)
    vel = np.vstack([vel, np.ones((9, 4))])
    cfg, keep, rec = O.make_cfg(L, rng=O.RNG_PHILOX, seed=1, step=0, walls=2, rho=10, record=True)
    O.collide(cfg, a, pos, vel)
    assert np.isnan(rec[2][3])


def test_mt_draw_order_with_walls():
    """Per cell with num > 0: [3 normals if a deficient wall cell] ->
    [gamma if rel > 0 and num > 1] -> 2 uniforms of the axis."""
    L = (3, 3, 4)
    rng = np.random.default_rng(5)
    n = 200
    pos = rng.uniform(0, 1, (n, 3)) * np.array([L[0], L[1], L[2] - 1.0])
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    a = np.array([0.2, 0.7, 0.1])
    mt = O.MTRand(321)
    cfg, keep, rec = O.make_cfg(L, rng=O.RNG_MT, mt=mt, walls=1, rho=10, record=True)
    O.collide(cfg, a, pos, vel)
    replay = O.MTRand(321)
    counts = np.bincount(O.box(L, a, pos), minlength=int(np.prod(L)))
    seen_virt = 0
    for c in range(int(np.prod(L))):
        if counts[c] == 0:
            assert np.isnan(rec[0][c])
            continue
        num = counts[c]
        if _wall_layer(L, c) and num < 10:
            nr = [replay.normal() for _ in range(3)]
            np.testing.assert_array_equal(rec[2][3 * c:3 * c + 3], nr)
            num = 10
            seen_virt += 1
        if not np.isnan(rec[0][c]):
            assert rec[0][c] == replay.gamma(num)
        ax = replay.axis()
        np.testing.assert_array_equal(rec[1][3 * c:3 * c + 3], ax)
    assert seen_virt > 5
