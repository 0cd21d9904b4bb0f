"""Pins of the oracle's geometry, binning and collision.

* box() (mpcd.cpp:130-138): brute force with exact rational arithmetic
  (fractions) on random and on-face positions, power-of-two and other boxes.
* Bin order: numpy's stable argsort of the cell ids over insertion order.
* Rotation matrix: with the exact cos/sin of 130 degrees it equals scipy's
  rotation by 130 degrees about the axis; the reference's truncated
  constants stay within 1e-7 of it.
* Collision (mpcd.cpp:140-247) with injected draws: per-cell momentum
  conserved; thermostat sets the relative kinetic energy to the injected
  Gamma variate (exact rotation); without thermostat and with exact
  constants the kinetic energy of every cell is conserved; the component of
  the relative velocity along the axis is untouched; a planted sign error
  (negated sin) is caught by the scipy comparison.
* Fluid::create (fluid.cpp:55-98): draw order against numpy's MT19937,
  zero mean velocity, kT = 1.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle as O
from cases import faithful_case, random_draws


def box_exact(L, a, x):
    out = []
    for k in range(3):
        s = Fraction(float(x[k]) + float(a[k]))  # the FP64 sum, as mpcd.cpp:131 forms it
        f = s.numerator // s.denominator
        out.append(f % L[k])
    return out[0] + L[0] * (out[1] + L[1] * out[2])


@pytest.mark.parametrize("L", [(32, 32, 32), (12, 10, 9), (100, 7, 3)])
def test_box_brute_force(L):
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 1, 3)
    x = rng.uniform(-3 * np.array(L), 4 * np.array(L), (3000, 3))
    # positions exactly on shifted cell faces
    k = rng.integers(-2 * L[0], 2 * L[0], (500, 3)).astype(np.float64)
    x = np.vstack([x, k - a])
    got = O.box(L, a, x)
    ref = [box_exact(L, a, xi) for xi in x]
    assert np.array_equal(got, ref)


def test_bin_is_stable_counting_sort():
    rng = np.random.default_rng(2)
    cell = rng.integers(0, 50, 2000).astype(np.int32)
    start, perm, rank = O.bin_cells(cell, 50)
    assert np.array_equal(perm, np.argsort(cell, kind="stable"))
    assert np.array_equal(start[1:] - start[:-1], np.bincount(cell, minlength=50))
    for i in range(0, 2000, 97):
        assert perm[start[cell[i]] + rank[i]] == i


def test_rotmat_against_scipy():
    rng = np.random.default_rng(3)
    alpha = np.deg2rad(130.0)
    for _ in range(50):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        R = O.rotmat(ax, np.cos(alpha), np.sin(alpha))
        Rs = Rotation.from_rotvec(alpha * ax).as_matrix()
        assert np.allclose(R, Rs, atol=1e-14)
        Rt = O.rotmat(ax)  # reference's truncated constants
        assert np.allclose(Rt, Rs, atol=1e-7)
        # a planted sign mistake is caught by this comparison
        Rbad = O.rotmat(ax, np.cos(alpha), -np.sin(alpha))
        assert not np.allclose(Rbad, Rs, atol=1e-3)


def _cell_sums(L, a, pos, vel):
    cell = O.box(L, a, pos)
    n = int(np.prod(L))
    m = vel[:, 3]
    P = np.stack([np.bincount(cell, weights=vel[:, k] * m, minlength=n) for k in range(3)], 1)
    M = np.bincount(cell, weights=m, minlength=n)
    return cell, P, M


def test_collision_invariants():
    L = (10, 9, 8)
    rng = np.random.default_rng(4)
    case = faithful_case(L, seed=11, steps=0)
    pos, vel = case["pos0"], case["vel0"].copy()
    a = case["a0"]
    ncells = int(np.prod(L))
    g, ax = random_draws(ncells, rng)
    alpha = np.deg2rad(130.0)
    cell, P0, M = _cell_sums(L, a, pos, vel)
    # thermostat on, exact rotation: relative KE of every cell becomes G
    v = vel.copy()
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax, calpha=np.cos(alpha), salpha=np.sin(alpha))
    O.collide(cfg, a, pos, v)
    _, P1, _ = _cell_sums(L, a, pos, v)
    assert np.allclose(P1, P0, atol=1e-12)
    u = P1 / np.maximum(M, 1)[:, None]
    rel = np.bincount(cell, weights=((v[:, :3] - u[cell]) ** 2).sum(1), minlength=ncells)
    cnt = np.bincount(cell, minlength=ncells)
    sel = cnt > 1
    assert np.allclose(0.5 * rel[sel], g[sel], rtol=1e-11)
    # thermostat off, exact constants: KE per cell conserved, axial component kept
    v = vel.copy()
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax, thermostat=0,
                              calpha=np.cos(alpha), salpha=np.sin(alpha))
    O.collide(cfg, a, pos, v)
    e0 = np.bincount(cell, weights=(vel[:, :3] ** 2).sum(1), minlength=ncells)
    e1 = np.bincount(cell, weights=(v[:, :3] ** 2).sum(1), minlength=ncells)
    assert np.allclose(e1, e0, rtol=1e-12)
    axc = ax.reshape(-1, 3)[cell]
    u0 = (P0 / np.maximum(M, 1)[:, None])[cell]
    assert np.allclose(((vel[:, :3] - u0) * axc).sum(1), ((v[:, :3] - u0) * axc).sum(1), atol=1e-12)
    # rotation by 130 degrees of the relative velocity
    r0, r1 = vel[:, :3] - u0, v[:, :3] - u0
    rp0 = r0 - (r0 * axc).sum(1)[:, None] * axc
    rp1 = r1 - (r1 * axc).sum(1)[:, None] * axc
    ok = (np.linalg.norm(rp0, axis=1) > 1e-6) & (cnt[cell] > 1)
    cosang = (rp0 * rp1).sum(1)[ok] / (np.linalg.norm(rp0, axis=1) * np.linalg.norm(rp1, axis=1))[ok]
    assert np.allclose(cosang, np.cos(alpha), atol=1e-9)


def test_collision_with_beads_conserves_momentum():
    L = (8, 8, 8)
    rng = np.random.default_rng(6)
    case = faithful_case(L, seed=12, steps=0)
    pos, vel = case["pos0"], case["vel0"].copy()
    nb = 200
    bpos = rng.uniform(0, 8, (nb, 3))
    bvel = np.zeros((nb, 4))
    bvel[:, :3] = rng.normal(size=(nb, 3))
    bvel[:, 3] = 5.0
    g, ax = random_draws(512, rng)
    p0 = vel[:, :3].sum(0) + (bvel[:, :3] * 5).sum(0)
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
    O.collide(cfg, case["a0"], pos, vel, bpos, bvel)
    p1 = vel[:, :3].sum(0) + (bvel[:, :3] * 5).sum(0)
    assert np.allclose(p0, p1, atol=1e-10)


def test_fluid_create_draw_order():
    L = (4, 5, 6)
    n = 50
    mt = O.MTRand(5489)
    pos, vel = O.fluid_create(mt, L, n)
    raw = np_raw = np.random.MT19937(0)
    np_raw._legacy_seeding(5489)
    r = raw.random_raw(9 * n)
    u = (r + 0.5) / 4294967296.0
    x0 = u[0:3] * np.array(L)
    assert np.array_equal(pos[0], x0)
    vx = np.sqrt(-2 * np.log(u[3])) * np.cos(2 * np.pi * u[4])
    # velocities have the mean removed afterwards; compare the difference pattern
    vx_all = []
    for i in range(n):
        b = 9 * i
        vx_all.append(np.sqrt(-2 * np.log(u[b + 3])) * np.cos(2 * np.pi * u[b + 4]))
    vx_all = np.array(vx_all)
    assert np.allclose(vel[:, 0], vx_all - vx_all.mean(), atol=1e-13)
    assert abs(vx_all[0] - vx) < 1e-15
    assert np.all(vel[:, 3] == 1.0)


def test_fluid_create_statistics():
    L = (20, 20, 20)
    mt = O.MTRand(9)
    pos, vel = O.fluid_create(mt, L, 80000)
    assert np.allclose(vel[:, :3].sum(0), 0, atol=1e-9)
    assert abs((vel[:, :3] ** 2).mean() - 1.0) < 0.01
    assert np.all((pos >= 0) & (pos < 20))
