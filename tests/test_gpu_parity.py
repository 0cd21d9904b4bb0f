"""GPU parity: the gfx950 step against the CPU oracle on identical states.

FP64 form: fed the reference's own per-cell draws (recorded from the
oracle's single-thread MT19937 run), every cell id, every particle's place
in its cell's order, every position and every velocity must come out
bit-identical to the oracle (both sides round each FP64 operation
separately; tolerance 0).  FP32-bulk form: bit-identical to the oracle's
restatement of that form.  Philox draws (no injection): within 1e-12
relative, the libm difference between glibc and the GPU math library.
"""
import numpy as np
import pytest

import oracle as O
from cases import faithful_case, periodic_diff, random_draws, rel_err

pytestmark = pytest.mark.gpu

H = 0.1


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("L", [(12, 10, 9), (32, 32, 32)])
def test_fp64_faithful_bit_exact(mcm, L):
    steps = 3 if L[0] < 32 else 2
    _print_case("fp64_faithful", L=L, steps=steps)
    case = faithful_case(L, seed=5489 + L[0], steps=steps, h=H)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=7) as m:
        m.set_shift(case["a0"])
        m.upload_fluid(case["pos0"], case["vel0"])
        m.arm_debug(True)
        ncells = int(np.prod(L))
        for s, st in enumerate(case["steps"]):
            m.step(H, next_shift=st["a_next"], inj_gamma=st["gamma"], inj_axis=st["axis"])
            cell, rank = m.debug_binning()
            ref_cell = O.box(L, st["a"], st["pos_before"])
            _, _, ref_rank = O.bin_cells(ref_cell, ncells)
            assert np.array_equal(cell, ref_cell), f"step {s}: cell ids differ"
            assert np.array_equal(rank, ref_rank), f"step {s}: per-cell order differs"
            pos, vel = m.download_fluid()
            assert np.array_equal(pos, st["pos"]), f"step {s}: positions differ, max {np.abs(pos - st['pos']).max()}"
            assert np.array_equal(vel[:, :3], st["vel"][:, :3]), \
                f"step {s}: velocities differ, max rel {rel_err(vel[:, :3], st['vel'][:, :3])}"
            d = m.diagnostics()
            assert d["flags"] == 0
            assert d["n"] == len(pos)
            assert abs(d["diss_last"] - st["diss"]) <= 1e-9 * max(1.0, abs(st["diss"])) + 1e-9


@pytest.mark.parametrize("variant", [0, 1])
def test_fp64_without_ids_bit_exact(mcm, variant):
    """Without particle ids the FP64 form keeps each cell in the order of
    the bits of vx (ties by arrival).  The oracle keeps a cell in particle
    index order, so before each step its arrays are permuted into that
    order; then every position and velocity must come out bit-identical,
    compared as multisets (the device's storage order is free)."""
    L = (16, 12, 8)
    _print_case("fp64_without_ids", L=L, variant=variant)
    rng = np.random.default_rng(37 + variant)
    case = faithful_case(L, seed=4242, steps=0)
    pos, vel = case["pos0"].copy(), case["vel0"].copy()
    a = case["a0"]
    ncells = int(np.prod(L))
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=False, seed=3, tile_variant=variant) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        for s in range(3):
            order = np.argsort(np.ascontiguousarray(vel[:, 0]).view(np.uint64), kind="stable")
            pos, vel = np.ascontiguousarray(pos[order]), np.ascontiguousarray(vel[order])
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            O.collide(cfg, a, pos, vel)
            O.stream(pos, vel, H)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            gp, gv = m.download_fluid()
            got = np.hstack([gp, gv[:, :3]])
            ref = np.hstack([pos, vel[:, :3]])
            got = got[np.lexsort(got.T[::-1])]
            ref = ref[np.lexsort(ref.T[::-1])]
            assert np.array_equal(got, ref), f"step {s} differs, max {np.abs(got - ref).max()}"
            assert m.diagnostics()["flags"] == 0
            a = a_next


@pytest.mark.parametrize("L", [(12, 10, 9), (32, 32, 32)])
def test_fp32_bulk_matches_restatement(mcm, L):
    steps = 3
    _print_case("fp32_restatement", L=L, steps=steps)
    rng = np.random.default_rng(11)
    case = faithful_case(L, seed=777, steps=0)
    a = case["a0"]
    ncells = int(np.prod(L))
    cell, off, w = O.f32_encode(L, a, case["pos0"], case["vel0"])
    with mcm.MPCD(L, precision=mcm.FP32BULK, track_ids=True, seed=3) as m:
        m.set_shift(a)
        m.upload_fluid(case["pos0"], case["vel0"])
        m.arm_debug(True)
        q0 = None
        for s in range(steps):
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0.0, 1.0, 3)
            cell_before = cell.copy()
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            O.f32_step(cfg, a, a_next, H, cell, off, w, ids=np.arange(len(cell)))
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            dcell, _ = m.debug_binning()
            assert np.array_equal(dcell, cell_before), f"step {s}: cell ids differ"
            pos, vel = m.download_fluid()
            ref_pos, ref_vel = O.f32_decode(L, a_next, cell, off, w)
            assert np.array_equal(vel[:, :3], ref_vel[:, :3]), f"step {s}: velocities differ"
            assert np.array_equal(pos, ref_pos), f"step {s}: positions differ"
            d = m.diagnostics()
            assert d["flags"] == 0
            q = d["momentum_q"]
            if q0 is None:
                q0 = np.round(w.astype(np.float64).sum(axis=0) * 2.0 ** 21).astype(np.int64)
            assert np.array_equal(q, q0), "FP32-bulk momentum must be conserved exactly"
            a = a_next


@pytest.mark.parametrize("track", [True, False])
@pytest.mark.parametrize("mode", ["coarse", "uniform"])
def test_fp32_owner_ties(mcm, track, mode):
    """Velocities that tie: many cells hold particles with equal velocity
    bits, so the momentum residual's owner is decided by the later key
    stages, by the id, or (identical records) by any claim.  "coarse":
    velocities on a 0.5 lattice plus duplicated particles; "uniform": every
    particle has the same velocity (it keeps it: u = v in every cell), so
    every particle of a cell is a candidate, more than the kernel's queue
    holds."""
    L = (12, 10, 9)
    _print_case("fp32_owner_ties", L=L, track=track, mode=mode)
    rng = np.random.default_rng(23)
    ncells = int(np.prod(L))
    n = 10 * ncells
    pos0 = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel0 = np.ones((n, 4))
    if mode == "coarse":
        vel0[:, :3] = rng.integers(-3, 4, (n, 3)) * 0.5
        dup, src = rng.choice(n, n // 10, replace=False), rng.choice(n, n // 10)
        pos0[dup], vel0[dup] = pos0[src], vel0[src]
    else:
        vel0[:, :3] = [0.5, -0.25, 0.75]
    a = rng.uniform(0, 1, 3)
    cell, off, w = O.f32_encode(L, a, pos0, vel0)
    q0 = np.round(w.astype(np.float64).sum(axis=0) * 2.0 ** 21).astype(np.int64)
    with mcm.MPCD(L, precision=mcm.FP32BULK, track_ids=track) as m:
        m.set_shift(a)
        m.upload_fluid(pos0, vel0)
        for s in range(3):
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            O.f32_step(cfg, a, a_next, H, cell, off, w, ids=np.arange(n) if track else None)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            pos, vel = m.download_fluid()
            rp, rv = O.f32_decode(L, a_next, cell, off, w)
            if track:
                assert np.array_equal(pos, rp) and np.array_equal(vel, rv), f"step {s} differs"
            else:  # storage order is free: compare the records as multisets
                got = np.hstack([pos, vel[:, :3]])
                ref = np.hstack([rp, rv[:, :3]])
                got = got[np.lexsort(got.T[::-1])]
                ref = ref[np.lexsort(ref.T[::-1])]
                assert np.array_equal(got, ref), f"step {s} differs"
            d = m.diagnostics()
            assert d["flags"] == 0
            assert np.array_equal(d["momentum_q"], q0)
            a = a_next


def test_fp64_philox_close_to_oracle(mcm):
    L = (16, 12, 8)
    _print_case("fp64_philox", L=L)
    case = faithful_case(L, seed=99, steps=0)
    seed = 1234
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=seed) as m:
        m.set_shift(case["a0"])
        m.upload_fluid(case["pos0"], case["vel0"])
        pos, vel = case["pos0"].copy(), case["vel0"].copy()
        a = case["a0"]
        for s in range(2):
            a_next = O.philox_shift(seed, s + 1)
            assert np.array_equal(m.draw_shift(), a_next)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_PHILOX, seed=seed, step=s)
            O.collide(cfg, a, pos, vel)
            O.stream(pos, vel, H)
            m.step(H)
            a = a_next
        gp, gv = m.download_fluid()
        assert rel_err(gv[:, :3], vel[:, :3]) < 1e-9
        assert np.max(np.abs(gp - pos)) < 1e-10
        assert np.allclose(m.get_shift(), a)


def test_beads_join_collision(mcm):
    L = (12, 10, 9)
    _print_case("beads", L=L)
    rng = np.random.default_rng(5)
    case = faithful_case(L, seed=31, steps=0)
    a = case["a0"]
    ncells = int(np.prod(L))
    nb = 400
    bpos = rng.uniform(0, 1, (nb, 3)) * np.array(L) + rng.normal(0, 3, (nb, 3))
    bvel = np.zeros((nb, 4))
    bvel[:, :3] = rng.normal(0, 0.3, (nb, 3))
    bvel[:, 3] = 5.0  # MDObject::partmass (mdobject.h:76)
    g, ax = random_draws(ncells, rng)
    a_next = rng.uniform(0, 1, 3)
    # FP64
    pos, vel = case["pos0"].copy(), case["vel0"].copy()
    rb = bvel.copy()
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
    O.collide(cfg, a, pos, vel, bpos, rb)
    O.stream(pos, vel, H)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(case["pos0"], case["vel0"])
        m.stage_beads(bpos, bvel)
        m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
        gb = m.fetch_beads(nb)
        gp, gv = m.download_fluid()
    assert np.array_equal(gb, rb), f"bead velocities differ, max {np.abs(gb - rb).max()}"
    assert np.array_equal(gv[:, :3], vel[:, :3])
    assert np.array_equal(gp, pos)
    # FP32-bulk
    cell, off, w = O.f32_encode(L, a, case["pos0"], case["vel0"])
    rb32 = bvel.copy()
    cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
    O.f32_step(cfg, a, a_next, H, cell, off, w, bpos, rb32, ids=np.arange(len(cell)))
    with mcm.MPCD(L, precision=mcm.FP32BULK, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(case["pos0"], case["vel0"])
        m.stage_beads(bpos, bvel)
        m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
        gb = m.fetch_beads(nb)
        gp, gv = m.download_fluid()
    ref_pos, ref_vel = O.f32_decode(L, a_next, cell, off, w)
    assert np.array_equal(gb, rb32)
    assert np.array_equal(gv, ref_vel)
    assert np.array_equal(gp, ref_pos)


def test_fp32_vs_fp64_one_step(mcm):
    """FP32-bulk against FP64 on the same injected state: 1e-5 relative."""
    L = (16, 16, 16)
    _print_case("fp32_vs_fp64", L=L)
    rng = np.random.default_rng(17)
    case = faithful_case(L, seed=4242, steps=0)
    ncells = int(np.prod(L))
    g, ax = random_draws(ncells, rng)
    a_next = rng.uniform(0, 1, 3)
    out = {}
    for prec in (mcm.FP64, mcm.FP32BULK):
        with mcm.MPCD(L, precision=prec, track_ids=True) as m:
            m.set_shift(case["a0"])
            m.upload_fluid(case["pos0"], case["vel0"])
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            out[prec] = m.download_fluid()
    p64, v64 = out[mcm.FP64]
    p32, v32 = out[mcm.FP32BULK]
    scale = np.maximum(np.linalg.norm(v64[:, :3], axis=1), 1.0)
    err = np.linalg.norm(v32[:, :3] - v64[:, :3], axis=1) / scale
    # cells can differ for particles within ~1e-7 of a face; those few collide elsewhere
    assert np.quantile(err, 0.999) < 1e-5
    assert np.max(np.abs(periodic_diff(p32, p64, L))) < 1e-3


def test_edge_sizes(mcm):
    """Single cell, two cells, empty fluid, and a large-shift move."""
    for L in [(1, 1, 1), (2, 1, 3), (9, 17, 5)]:
        _print_case("edge", L=L)
        case = faithful_case(L, seed=3 + L[1], steps=2, h=H)
        with mcm.MPCD(L, precision=mcm.FP64, track_ids=True) as m:
            m.set_shift(case["a0"])
            m.upload_fluid(case["pos0"], case["vel0"])
            for st in case["steps"]:
                m.step(H, next_shift=st["a_next"], inj_gamma=st["gamma"], inj_axis=st["axis"])
            pos, vel = m.download_fluid()
            assert np.array_equal(pos, case["steps"][-1]["pos"])
            assert np.array_equal(vel[:, :3], case["steps"][-1]["vel"][:, :3])
    with mcm.MPCD((8, 8, 8), precision=mcm.FP32BULK) as m:
        m.upload_fluid(np.zeros((0, 3)), np.zeros((0, 4)))
        m.step(H)
        assert m.download_fluid()[0].shape == (0, 3)
        assert m.diagnostics()["flags"] == 0


@pytest.mark.parametrize("prec_name", ["FP64", "FP32BULK"])
def test_every_tile_variant_same_result(mcm, prec_name):
    """Every tile shape of the step kernel gives the identical result."""
    prec = getattr(mcm, prec_name)
    L = (20, 12, 9)
    case = faithful_case(L, seed=1701, steps=2, h=H)
    outs = []
    for v, shape in enumerate(mcm.tile_variants(prec)):
        _print_case("variant", prec=prec_name, variant=v, shape=shape)
        with mcm.MPCD(L, precision=prec, track_ids=True, tile_variant=v) as m:
            m.set_shift(case["a0"])
            m.upload_fluid(case["pos0"], case["vel0"])
            for st in case["steps"]:
                m.step(H, next_shift=st["a_next"], inj_gamma=st["gamma"], inj_axis=st["axis"])
            outs.append(m.download_fluid())
            assert m.diagnostics()["flags"] == 0
    for p, v in outs[1:]:
        assert np.array_equal(p, outs[0][0]) and np.array_equal(v, outs[0][1])
    if prec == mcm.FP64:
        assert np.array_equal(outs[0][0], case["steps"][-1]["pos"])
