"""GPU parity where the fluid is locally dense: tiles above the step
kernel's particle capacity.

Such tiles are deferred by the step kernel to a second pass that holds
more particles per tile (the same tile code).  The result must be the
oracle's, bit for bit, as for any other tile (FP64 form with the injected
per-cell draws; FP32-bulk form against its restatement), with the deferral
visible in the diagnostics.  A tile beyond even the second pass's capacity
is reported.  An ensemble with a dense box matches the
box stepped alone.
"""
import numpy as np
import pytest

import oracle as O
from cases import random_draws

pytestmark = pytest.mark.gpu

H = 0.1
L = (16, 16, 16)


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _dense_fluid(rng, extra, lo=8.0, width=3.0):
    """10 particles per cell everywhere plus `extra` in a cube of `width`
    cells; [8, 11) plus any shift in (0, 1) lies in cells 8..11, inside one
    tile of every tile shape."""
    n = 10 * int(np.prod(L))
    pos = np.vstack([rng.uniform(0, 1, (n, 3)) * np.array(L), lo + width * rng.uniform(0, 1, (extra, 3))])
    vel = np.ones((len(pos), 4))
    vel[:, :3] = rng.normal(size=(len(pos), 3))
    return pos, vel


# extra particles in the dense cube that put its tile between the step
# kernel's capacity and the large pass's: the FP64 form's default 4x4x4 tile
# holds 1,024 (home segments of 2,048), the FP32-bulk form's 8x4x4 one 1,792
# (home segments of 4,096)
EXTRA = {"FP64": 800, "FP32BULK": 1600}


def test_fp64_dense_tiles_bit_exact(mcm):
    _print_case("fp64_dense", L=L)
    rng = np.random.default_rng(91)
    pos, vel = _dense_fluid(rng, EXTRA["FP64"])
    ncells = int(np.prod(L))
    a = rng.uniform(0, 1, 3)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(3):
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            O.collide(cfg, a, pos, vel)
            O.stream(pos, vel, H)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            gp, gv = m.download_fluid()
            d = m.diagnostics()
            assert d["flags"] == 0
            assert d["deferred_tiles"] >= 1 and d["max_tile_load"] > d["tile_capacity"], d
            assert np.array_equal(gp, pos), f"step {s}: positions differ, max {np.abs(gp - pos).max()}"
            assert np.array_equal(gv[:, :3], vel[:, :3]), f"step {s}: velocities differ"
            a = a_next


def test_fp32_dense_tiles_match_restatement(mcm):
    _print_case("fp32_dense", L=L)
    rng = np.random.default_rng(92)
    pos, vel = _dense_fluid(rng, EXTRA["FP32BULK"])
    ncells = int(np.prod(L))
    a = rng.uniform(0, 1, 3)
    cell, off, w = O.f32_encode(L, a, pos, vel)
    with mcm.MPCD(L, precision=mcm.FP32BULK, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(3):
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            O.f32_step(cfg, a, a_next, H, cell, off, w, ids=np.arange(len(cell)))
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            gp, gv = m.download_fluid()
            rp, rv = O.f32_decode(L, a_next, cell, off, w)
            d = m.diagnostics()
            assert d["flags"] == 0
            assert d["deferred_tiles"] >= 1 and d["max_tile_load"] > d["tile_capacity"], d
            assert np.array_equal(gv[:, :3], rv[:, :3]), f"step {s}: velocities differ"
            assert np.array_equal(gp, rp), f"step {s}: positions differ"
            a = a_next


@pytest.mark.parametrize("prec_name", ["FP64", "FP32BULK"])
def test_beyond_second_capacity_is_reported(mcm, prec_name):
    """A tile of more than 32 particles per cell is reported, not silently
    stepped: its home segment (sized for the deferred pass) is full at the
    upload."""
    _print_case("dense_reported", prec=prec_name)
    rng = np.random.default_rng(93)
    pos, vel = _dense_fluid(rng, 6000, lo=8.2, width=1.5)
    with mcm.MPCD(L, precision=getattr(mcm, prec_name)) as m:
        m.set_shift(np.array([0.5, 0.5, 0.5]))
        with pytest.raises(mcm.MPCDError, match="home segment full"):
            m.upload_fluid(pos, vel)


@pytest.mark.parametrize("prec_name", ["FP64", "FP32BULK"])
def test_ensemble_with_dense_box(mcm, prec_name):
    _print_case("ensemble_dense", prec=prec_name)
    prec = getattr(mcm, prec_name)
    rng = np.random.default_rng(94)
    pos, vel = _dense_fluid(rng, EXTRA[prec_name])

    def make(dense):
        m = mcm.MPCD(L, precision=prec, track_ids=True, seed=5 if dense else 6)
        if dense:
            m.upload_fluid(pos, vel)
        else:
            m.create_fluid(10 * int(np.prod(L)))
        return m

    alone = [make(True), make(False)]
    ens = [make(True), make(False)]
    try:
        for m in alone:
            for _ in range(3):
                m.step(H)
        with mcm.Ensemble(ens) as e:
            e.step(H, 3)
            e.sync()
            for x, y in zip(ens, alone):
                px, vx = x.download_fluid()
                py, vy = y.download_fluid()
                assert np.array_equal(px, py) and np.array_equal(vx, vy)
            dx = ens[0].diagnostics()
            assert dx["flags"] == 0 and dx["deferred_tiles"] >= 1
    finally:
        for m in alone + ens:
            m.close()


@pytest.mark.parametrize("prec_name,rule_name", [("FP64", "AT"), ("FP64", "AT_A"), ("FP32BULK", "AT"),
                                                 ("FP32BULK", "AT_A")])
def test_at_dense_tiles_bit_exact(mcm, prec_name, rule_name):
    """MPC-AT in tiles above the step kernel's capacity: the large pass's
    instantiations of the rule, bit for bit with the oracle (injected
    normals by id)."""
    _print_case("at_dense", prec=prec_name, rule=rule_name, L=L)
    rng = np.random.default_rng(95)
    # MPC-AT runs on the 4x4x4 tile (1,024 particles, home segments of 2,048):
    # 800 more in the dense cube put its tile between the two
    pos, vel = _dense_fluid(rng, 800)
    n = len(pos)
    rule = getattr(O, rule_name)
    a = rng.uniform(0, 1, 3)
    f32 = prec_name == "FP32BULK"
    if f32:
        cell, off, w = O.f32_encode(L, a, pos, vel)
    with mcm.MPCD(L, precision=getattr(mcm, prec_name), track_ids=True, collision=getattr(mcm, rule_name)) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(2):
            ran = rng.normal(size=3 * n)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, rule=rule, ran=ran)
            if f32:
                O.f32_step(cfg, a, a_next, H, cell, off, w, ids=np.arange(n, dtype=np.int32))
                rp, rv = O.f32_decode(L, a_next, cell, off, w)
            else:
                O.collide(cfg, a, pos, vel)
                O.stream(pos, vel, H)
                rp, rv = pos, vel
            m.step(H, next_shift=a_next, inj_ran=ran)
            gp, gv = m.download_fluid()
            d = m.diagnostics()
            assert d["flags"] == 0
            assert d["deferred_tiles"] >= 1, d
            assert np.array_equal(gv[:, :3], rv[:, :3]), f"step {s}: velocities differ, max {np.abs(gv - rv).max()}"
            assert np.array_equal(gp, rp), f"step {s}: positions differ"
            a = a_next
