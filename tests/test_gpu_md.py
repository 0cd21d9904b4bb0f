"""GPU parity of the bead MD half of the step (sub-steps, forces, beat).

The device engine evaluates each bead's force from an ordered term list
built from the swimmer description; the oracle runs the reference's own
object loops (md_oracle.c).  Both round every FP64 operation separately,
so positions, velocities, rest lengths, the beat state and globalt must be
bit-identical after every MPCD step -- alone, and with the beads taking
part in the solvent collision (FP64 form and FP32-bulk form, injected
per-cell draws).
"""
import numpy as np
import pytest

import md_cases as M
import oracle as O
from cases import random_draws

pytestmark = pytest.mark.gpu

H = 0.05
DT = 0.001


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _compare(st, o, step):
    assert np.array_equal(st["pos"], o.pos), f"step {step}: bead positions differ, max {np.abs(st['pos'] - o.pos).max()}"
    assert np.array_equal(st["vel"][:, :3], o.vel[:, :3]), \
        f"step {step}: bead velocities differ, max {np.abs(st['vel'][:, :3] - o.vel[:, :3]).max()}"
    assert np.array_equal(st["newlength"], np.concatenate(o.newlength))
    assert np.array_equal(st["oldlength"], np.concatenate(o.oldlength))
    assert [tuple(b) for b in st["beat"]] == o.cilia_state()
    assert st["globalt"] == o.globalt


@pytest.mark.parametrize("constant_beat,anchor", [(True, False), (False, True)])
def test_md_substeps_bit_exact(mcm, constant_beat, anchor):
    _print_case("md_substeps", constant_beat=constant_beat, anchor=anchor)
    sw = M.swimmer(seed=11, constant_beat=constant_beat, mobile_anchor=anchor)
    o = O.MDOracle(sw, globalt=0.0, ststep=DT, tstep=H)
    with mcm.MPCD((16, 16, 16), precision=mcm.FP64) as m:
        m.md_attach(sw, globalt=0.0, ststep=DT)
        for s in range(6):
            m.md_run(H)
            o.step()
            _compare(m.md_fetch(), o, s)
    # the beat went through a stroke transition in these steps
    if constant_beat:
        assert any(mid != 2.0 for mid, _ in o.cilia_state())


@pytest.mark.parametrize("tstep,ststep,steps", [(0.025, 0.001, 8), (0.05, 0.0013, 22)])
def test_md_refresh_schedules_that_change_every_step(mcm, tstep, ststep, steps):
    """Step lengths whose rest-length refresh schedules differ from step to
    step: two alternating ones (0.025 in sub-steps of 0.001), and more
    distinct ones than the engine keeps graphs for (0.05 in 0.0013, 20 in
    22 steps), the rest served by the graph that refreshes at every
    evaluation; every step bit-identical to the oracle."""
    _print_case("md_schedules", tstep=tstep, ststep=ststep, steps=steps)
    sw = M.swimmer(seed=11, constant_beat=False, mobile_anchor=False)
    o = O.MDOracle(sw, globalt=0.0, ststep=ststep, tstep=tstep)
    with mcm.MPCD((16, 16, 16), precision=mcm.FP64) as m:
        m.md_attach(sw, globalt=0.0, ststep=ststep)
        for s in range(steps):
            m.md_run(tstep)
            o.step()
            _compare(m.md_fetch(), o, s)


@pytest.mark.parametrize("prec_name,anchor", [("FP64", False), ("FP32BULK", False), ("FP64", True),
                                              ("FP32BULK", True)])
def test_md_with_collision(mcm, prec_name, anchor):
    """mpcd_step with a swimmer attached: collide (solvent + beads), stream,
    then the MD half -- against the oracle's collision followed by its MD.
    With anchor, bead 0 of every rod is fixed (startidx 1): fixed beads
    neither move nor join the collision (MDObject::init, md_loop_stop add
    only beads j >= startidx, src/libcilia/mdobject.cpp:62-72,153-160), so
    the oracle collides the mobile beads alone, in index order."""
    prec = getattr(mcm, prec_name)
    L = (16, 16, 16)
    _print_case("md_collision", prec=prec_name, L=L, anchor=anchor)
    rng = np.random.default_rng(12)
    ncells = int(np.prod(L))
    n = 10 * ncells
    pos0 = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel0 = np.ones((n, 4))
    vel0[:, :3] = rng.normal(0, 1, (n, 3))
    sw = M.swimmer(seed=13, mobile_anchor=anchor)
    o = O.MDOracle(sw, globalt=0.0, ststep=DT, tstep=H)
    mob = np.ones(len(o.pos), dtype=bool) if sw["mobile"] is None else sw["mobile"].astype(bool)
    assert anchor == (not mob.all())
    a = rng.uniform(0, 1, 3)
    ids = np.arange(n)
    if prec == mcm.FP64:
        pos, vel = pos0.copy(), vel0.copy()
    else:
        cell, off, w = O.f32_encode(L, a, pos0, vel0)
    with mcm.MPCD(L, precision=prec, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(pos0, vel0)
        m.md_attach(sw, globalt=0.0, ststep=DT)
        for s in range(3):
            g, ax = random_draws(ncells, rng)
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax)
            bp, bv = np.ascontiguousarray(o.pos[mob]), np.ascontiguousarray(o.vel[mob])
            if prec == mcm.FP64:
                O.collide(cfg, a, pos, vel, bp, bv)
                O.stream(pos, vel, H)
            else:
                O.f32_step(cfg, a, a_next, H, cell, off, w, bp, bv, ids=ids)
            o.vel[mob] = bv
            o.step()
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            _compare(m.md_fetch(), o, s)
            gp, gv = m.download_fluid()
            if prec == mcm.FP64:
                assert np.array_equal(gp, pos) and np.array_equal(gv[:, :3], vel[:, :3]), f"step {s}: solvent differs"
            else:
                rp, rv = O.f32_decode(L, a_next, cell, off, w)
                assert np.array_equal(gp, rp) and np.array_equal(gv, rv), f"step {s}: solvent differs"
            assert m.diagnostics()["flags"] == 0
            a = a_next


def test_md_wide_path_bit_exact(mcm):
    """A swimmer large enough for the one-launch-per-phase path (more than
    65536 force terms): 643-bead shell, 150 cilia of 3 x 20 beads."""
    from mcm_amd.swimmer import clami
    _print_case("md_wide", cilia=150)
    sw = clami(levels=3, radius=10.0, center=(24.0, 24.0, 24.0), n_cilia=150, rodlength=20, seed=3,
               power_stroke_duration=0.06, recovery_stroke_duration=0.1)
    o = O.MDOracle(sw, globalt=0.0, ststep=DT, tstep=H)
    with mcm.MPCD((48, 48, 48), precision=mcm.FP64) as m:
        m.md_attach(sw, globalt=0.0, ststep=DT)
        for s in range(3):
            m.md_run(H)
            o.step()
            _compare(m.md_fetch(), o, s)
    assert any(mid != 2.0 for mid, _ in o.cilia_state())
