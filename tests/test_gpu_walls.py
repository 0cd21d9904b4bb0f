"""GPU parity of the planar walls (the reference's SLIP / NOSLIP build):
virtual particles in the bottom and top cell layers and bounce-back of the
solvent and of the swimmer's beads.

FP64 form fed the reference's own MT19937 draws (per cell: 3 virtual
normals, gamma, axis, recorded by the oracle): positions, velocities and
the per-cell order bit-identical to the oracle, tolerance 0.  Philox draws
(no injection): within 1e-12 relative of the oracle's Philox stream, the
libm difference between glibc and the GPU math library.
"""
import numpy as np
import pytest

import md_cases as M
import oracle as O
from cases import rel_err, walls_case

pytestmark = pytest.mark.gpu

H = 0.1


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("walls,vw", [(1, (0.0, 0.0)), (1, (0.25, -0.1)), (2, (0.0, 0.0))])
def test_walls_fp64_bit_exact(mcm, walls, vw):
    L = (10, 9, 8)
    steps = 4
    _print_case("walls_fp64", walls=walls, vw=vw, L=L, steps=steps)
    case = walls_case(L, seed=4242 + walls, steps=steps, walls=walls, h=H, vw=vw)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=7, walls=walls, wall_v=vw) as m:
        m.set_shift(case["a0"])
        m.upload_fluid(case["pos0"], case["vel0"])
        m.arm_debug(True)
        ncells = int(np.prod(L))
        for s, st in enumerate(case["steps"]):
            assert st["nvirt"] > 20
            m.step(H, next_shift=st["a_next"], inj_gamma=st["gamma"], inj_axis=st["axis"], inj_virt=st["virt"])
            cell, rank = m.debug_binning()
            ref_cell = O.box(L, st["a"], st["pos_before"])
            _, _, ref_rank = O.bin_cells(ref_cell, ncells)
            assert np.array_equal(cell, ref_cell), f"step {s}: cell ids differ"
            assert np.array_equal(rank, ref_rank), f"step {s}: per-cell order differs"
            pos, vel = m.download_fluid()
            assert np.array_equal(pos, st["pos"]), f"step {s}: positions differ, max {np.abs(pos - st['pos']).max()}"
            assert np.array_equal(vel[:, :3], st["vel"][:, :3]), \
                f"step {s}: velocities differ, max rel {rel_err(vel[:, :3], st['vel'][:, :3])}"
            assert pos[:, 2].min() >= 0.0 and pos[:, 2].max() <= L[2] - 1.0
            d = m.diagnostics()
            assert d["flags"] == 0
            assert abs(d["diss_last"] - st["diss"]) <= 1e-9 * max(1.0, abs(st["diss"])) + 1e-9
            np.testing.assert_allclose(d["momentum"], vel[:, :3].sum(axis=0), rtol=0, atol=1e-8)


def test_walls_fp64_philox_close_to_oracle(mcm):
    L = (12, 10, 9)
    seed = 31
    _print_case("walls_philox", L=L)
    case = walls_case(L, seed=5, steps=0, walls=1)
    pos, vel = case["pos0"].copy(), case["vel0"].copy()
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=seed, walls=1) as m:
        a = np.array([0.37, 0.81, 0.12])
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(3):
            a_next = m.draw_shift()
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=s, walls=1, rho=10)
            O.collide(cfg, a, pos, vel)
            O.stream_walls(pos, vel, H, L[2], 1)
            m.step(H, next_shift=a_next)
            gpos, gvel = m.download_fluid()
            assert rel_err(gvel[:, :3], vel[:, :3]) < 1e-9 or np.abs(gvel[:, :3] - vel[:, :3]).max() < 1e-11
            assert np.abs(gpos - pos).max() < 1e-11
            a = a_next


def test_walls_refused_where_unsupported(mcm):
    with pytest.raises(mcm.MPCDError):
        mcm.MPCD((8, 8, 8), precision=mcm.FP32BULK, walls=1, tile_variant=1)
    with pytest.raises(mcm.MPCDError):
        mcm.MPCD((8, 8, 1), precision=mcm.FP64, walls=2)
    with mcm.MPCD((8, 8, 8), precision=mcm.FP64, walls=1) as m:
        m.create_fluid(10 * 8 * 8 * 7)
        g = np.ones(512)
        ax = np.tile([0.0, 0.0, 1.0], 512)
        with pytest.raises(mcm.MPCDError):  # injected draws without the virtual normals
            m.step(0.1, inj_gamma=g, inj_axis=ax)


@pytest.mark.parametrize("walls,side,vw", [(1, "bottom", (0.0, 0.0)), (1, "top", (0.5, -0.25)), (2, "top", (0.0, 0.0))])
def test_md_beads_bounce_at_walls(mcm, walls, side, vw):
    """The swimmer is driven into a wall: its beads bounce in the MD drift
    (MPCD::move) bit-identically to the oracle, and differently from a run
    without walls."""
    L = (16, 16, 24)
    _print_case("md_walls", walls=walls, side=side, vw=vw)
    sw = M.swimmer(seed=3, center=(8.0, 8.0, 10.0))
    if side == "bottom":
        sw["pos3"][:, 2] += 0.05 - sw["pos3"][:, 2].min()
        sw["vel4"][:, 2] -= 2.0
    else:
        sw["pos3"][:, 2] += (L[2] - 1.05) - sw["pos3"][:, 2].max()
        sw["vel4"][:, 2] += 2.0
    assert 0.0 < sw["pos3"][:, 2].min() and sw["pos3"][:, 2].max() < L[2] - 1
    o = O.MDOracle(sw, globalt=0.0, ststep=0.001, tstep=0.05, walls=walls, Lz=L[2], vw=vw)
    free = O.MDOracle(sw, globalt=0.0, ststep=0.001, tstep=0.05)
    with mcm.MPCD(L, precision=mcm.FP64, walls=walls, wall_v=vw) as m:
        m.md_attach(sw, globalt=0.0, ststep=0.001)
        for s in range(4):
            m.md_run(0.05)
            o.step()
            free.step()
            st = m.md_fetch()
            assert np.array_equal(st["pos"], o.pos), f"step {s}: max {np.abs(st['pos'] - o.pos).max()}"
            assert np.array_equal(st["vel"][:, :3], o.vel[:, :3])
        assert np.abs(o.pos - free.pos).max() > 0.1
