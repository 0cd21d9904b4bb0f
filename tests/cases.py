"""Shared state builders for the parity tests.

A "faithful" case reproduces the reference's own random stream with one
thread: MTRand(seed) draws the constructor's grid shift (mpcd.cpp:86-97),
then Fluid::create (fluid.cpp:55-98), then per step the collision's gamma
and axis draws in cell order (mpcd.cpp:140-247) and the next shift
(mpcd.cpp:249-253).  The oracle records the draws so the GPU can be fed the
exact same per-cell values.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def faithful_case(L, seed, steps, h=0.1, thermostat=True):
    mt = O.MTRand(seed)
    a = np.array([mt.dblexc() for _ in range(3)])
    n = 10 * int(np.prod(L))
    pos0, vel0 = O.fluid_create(mt, L, n)
    pos, vel = pos0.copy(), vel0.copy()
    steps_out = []
    for _ in range(steps):
        before = pos.copy()
        cfg, keep, rec = O.make_cfg(L, rng=O.RNG_MT, mt=mt, record=True, thermostat=int(thermostat))
        diss = O.collide(cfg, a, pos, vel)
        a_next = np.array([mt.dblexc() for _ in range(3)])
        O.stream(pos, vel, h)
        g = np.nan_to_num(rec[0], nan=1.0)
        ax = np.nan_to_num(rec[1], nan=0.0)
        steps_out.append(dict(a=a.copy(), a_next=a_next, gamma=g, axis=ax, diss=diss,
                              pos_before=before, pos=pos.copy(), vel=vel.copy(),
                              rel_violations=cfg.rel_violations))
        a = a_next
    return dict(L=tuple(L), a0=steps_out[0]["a"] if steps else a, pos0=pos0, vel0=vel0, steps=steps_out)


def random_draws(ncells, rng, nc_typ=10):
    """Per-cell injected gamma variates and unit axes."""
    g = rng.gamma(1.5 * (nc_typ - 1), 1.0, size=ncells)
    ax = rng.normal(size=(ncells, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    return g, ax.ravel()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def periodic_diff(p, q, L):
    d = np.asarray(p) - np.asarray(q)
    Ls = np.asarray(L, dtype=np.float64)
    return d - Ls * np.round(d / Ls)


def walls_case(L, seed, steps, walls, h=0.1, vw=(0.0, 0.0), rho=10):
    """Planar walls (walls 1 NOSLIP, 2 SLIP): the solvent fills z in
    [0, Lz - 1] (uniform positions, unit normal velocities from numpy), the
    per-step draws come from MTRand(seed) in the reference's order: per
    cell [3 virtual-particle normals] -> [gamma] -> axis, then the next
    shift.  The oracle records the draws for injection."""
    rng = np.random.default_rng(seed)
    mt = O.MTRand(seed)
    a = np.array([mt.dblexc() for _ in range(3)])
    n = rho * L[0] * L[1] * (L[2] - 1)
    pos0 = rng.uniform(0.0, 1.0, (n, 3)) * np.array([L[0], L[1], L[2] - 1.0])
    vel0 = np.ones((n, 4))
    vel0[:, :3] = rng.normal(size=(n, 3))
    vel0[:, :3] -= vel0[:, :3].mean(axis=0)
    pos, vel = pos0.copy(), vel0.copy()
    steps_out = []
    for _ in range(steps):
        before = pos.copy()
        cfg, keep, rec = O.make_cfg(L, rng=O.RNG_MT, mt=mt, record=True, walls=walls, rho=rho)
        diss = O.collide(cfg, a, pos, vel)
        a_next = np.array([mt.dblexc() for _ in range(3)])
        O.stream_walls(pos, vel, h, L[2], walls, vw)
        steps_out.append(dict(a=a.copy(), a_next=a_next, gamma=np.nan_to_num(rec[0], nan=1.0),
                              axis=np.nan_to_num(rec[1], nan=0.0), virt=np.nan_to_num(rec[2], nan=0.0),
                              nvirt=int(np.sum(~np.isnan(rec[2][0::3]))), diss=diss, pos_before=before,
                              pos=pos.copy(), vel=vel.copy()))
        a = a_next
    return dict(L=tuple(L), a0=steps_out[0]["a"] if steps else a, pos0=pos0, vel0=vel0, steps=steps_out)
