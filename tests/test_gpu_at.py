"""GPU parity of the MPC-AT collision rules (FP64 form).

With the MT19937 normals the oracle draws injected per particle id, every
cell id, per-cell rank, position and velocity must be bit-identical to the
oracle (tolerance 0), with and without beads.  With the Philox normals:
within 1e-12 of the oracle (the GPU's cospi against glibc's cos).  At a
large size: per-cell conservation shows in the totals, and the Andersen
thermostat holds kT.
"""
import numpy as np
import pytest

import oracle as O
from cases import rel_err

pytestmark = pytest.mark.gpu

H = 0.1


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _state(L, rng, nb=0):
    n = 10 * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    bpos = rng.uniform(0, 1, (nb, 3)) * np.array(L)
    bvel = np.full((nb, 4), 5.0)
    bvel[:, :3] = rng.normal(size=(nb, 3)) * 0.4
    return pos, vel, bpos, bvel


@pytest.mark.parametrize("rule_name", ["AT", "AT_A"])
def test_at_fp64_bit_exact(mcm, rule_name):
    L = (12, 10, 9)
    rule = getattr(O, rule_name)
    _print_case("at_fp64", rule=rule_name, L=L)
    rng = np.random.default_rng(31)
    pos, vel, _, _ = _state(L, rng)
    n = len(pos)
    ncells = int(np.prod(L))
    mt = O.MTRand(2024)
    a = rng.uniform(0, 1, 3)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, collision=getattr(mcm, rule_name)) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        m.arm_debug(True)
        for s in range(3):
            before = pos.copy()
            cfg, keep, rec = O.make_cfg(L, rng=O.RNG_MT, mt=mt, rule=rule, record=True, n_record_ran=n)
            diss = O.collide(cfg, a, pos, vel)
            a_next = rng.uniform(0, 1, 3)
            O.stream(pos, vel, H)
            m.step(H, next_shift=a_next, inj_ran=rec[3][:3 * n])
            cell, rank = m.debug_binning()
            ref_cell = O.box(L, a, before)
            _, _, ref_rank = O.bin_cells(ref_cell, ncells)
            assert np.array_equal(cell, ref_cell), f"step {s}: cell ids differ"
            assert np.array_equal(rank, ref_rank), f"step {s}: per-cell order differs"
            gp, gv = m.download_fluid()
            assert np.array_equal(gp, pos), f"step {s}: positions differ, max {np.abs(gp - pos).max()}"
            assert np.array_equal(gv[:, :3], vel[:, :3]), f"step {s}: velocities differ, max rel {rel_err(gv, vel)}"
            d = m.diagnostics()
            assert d["flags"] == 0
            assert abs(d["diss_last"] - diss) <= 1e-9 * max(1.0, abs(diss))
            a = a_next


def test_at_a_with_beads_bit_exact(mcm):
    L = (8, 8, 8)
    _print_case("at_beads", L=L)
    rng = np.random.default_rng(5)
    pos, vel, bpos, bvel = _state(L, rng, nb=90)
    n = len(pos)
    a = rng.uniform(0, 1, 3)
    ran = rng.normal(size=3 * (n + len(bpos)))
    bv = bvel.copy()
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, rule=O.AT_A, ran=ran)
    p1, v1 = pos.copy(), vel.copy()
    O.collide(cfg, a, p1, v1, bpos, bv)
    O.stream(p1, v1, H)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, collision=mcm.AT_A) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        m.stage_beads(bpos, bvel)
        m.step(H, next_shift=rng.uniform(0, 1, 3), inj_ran=ran)
        gp, gv = m.download_fluid()
        gb = m.fetch_beads(len(bpos))
    assert np.array_equal(gv[:, :3], v1[:, :3])
    assert np.array_equal(gp, p1)
    assert np.array_equal(gb, bv)


@pytest.mark.parametrize("rule_name", ["AT", "AT_A"])
def test_at_philox_close_to_oracle(mcm, rule_name):
    L = (12, 10, 9)
    rule = getattr(O, rule_name)
    _print_case("at_philox", rule=rule_name)
    rng = np.random.default_rng(8)
    pos, vel, _, _ = _state(L, rng)
    a = np.array([0.25, 0.5, 0.75])
    seed = 17
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=seed, collision=getattr(mcm, rule_name)) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(2):
            a_next = m.draw_shift()
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=s, rule=rule)
            O.collide(cfg, a, pos, vel)
            O.stream(pos, vel, H)
            m.step(H, next_shift=a_next)
            gp, gv = m.download_fluid()
            assert np.abs(gv[:, :3] - vel[:, :3]).max() < 1e-12
            assert np.abs(gp - pos).max() < 1e-12
            a = a_next


def test_at_a_large_box_invariants(mcm):
    L = (64, 64, 64)
    _print_case("at_large", L=L)
    n = 10 * int(np.prod(L))
    with mcm.MPCD(L, precision=mcm.FP64, seed=99, collision=mcm.AT_A) as m:
        m.create_fluid(n)
        for _ in range(10):
            m.step(H)
        d = m.diagnostics()
        assert d["flags"] == 0 and d["n"] == n
        assert np.max(np.abs(d["momentum"])) < 1e-6 * n ** 0.5
        assert abs(d["twice_kinetic"] / (3.0 * n) - 1.0) < 5e-3


def test_at_refused_where_unsupported(mcm):
    with pytest.raises(mcm.MPCDError):
        mcm.MPCD((8, 8, 8), precision=mcm.FP32BULK, collision=mcm.AT, tile_variant=2)
    with pytest.raises(mcm.MPCDError):
        mcm.MPCD((8, 8, 8), precision=mcm.FP64, collision=mcm.AT_A, walls=1)
    with pytest.raises(mcm.MPCDError):
        mcm.MPCD((8, 8, 8), precision=mcm.FP64, collision=mcm.AT, tile_variant=0)
