"""Every benchmark configuration at its own workload on the GPU.

C0  periodic 32^3, 327,680 solvent particles, FP64, SRD + thermostat,
    h = 0.1, a 643-bead shell of radius 5 with 16 cilia of 3 x 10 beads
    (1,123 beads), 50 MD sub-steps: 3 steps bit-identical to the oracle
    with the reference's own MT19937 draws injected (its Fluid::create,
    then per step the collision's draws in cell order and the next shift),
    the beads colliding with the solvent; then 1000 steps on the Philox
    stream: flags clean, total momentum (solvent + beads) to roundoff, kT.
C1  periodic 256^3, 167,772,160 particles, both forms, 20 steps: flags,
    particle count, momentum (FP32-bulk: the exact integer; FP64: roundoff),
    kT; one step of both forms from the same state, stored exactly in both,
    with the same injected draws (FP32-bulk within 1e-5 of FP64); the kinetic energy
    change without the thermostat against the truncated rotation constants'
    prediction.
C2  periodic 100^3, 10,000,000 particles, a 643-bead shell of radius 16
    with 400 cilia of 3 x 20 beads (24,643 beads), 50 MD sub-steps: one
    step of both forms from the same state with the same draws (beads,
    shell force and torque, solvent velocities within FP32 precision of
    FP64); 2000 steps per form: the swimmer's velocity agrees within the
    scatter of its block averages.
C3  64 independent swimmers in cubic boxes of edge 32..64 (86M particles),
    FP32-bulk, the tile variant chosen by the autotuner, 20 ensemble steps
    from the captured graph: every box's flags clean, and three sample
    boxes bit-identical to the same boxes stepped alone.
C4  periodic 128^3, 20,971,520 particles, MPC-AT+a, a 643-bead shell of
    radius 20 with 1000 cilia of 3 x 20 beads (60,643 beads), 100 MD
    sub-steps: one step of both forms from the same state with the same
    injected normals (beads, shell force and torque within 1e-5); 2000
    steps per form: total momentum drift, kT, and the last collision's
    angular momentum in the cells without beads (MPC-AT+a conserves it per
    cell).
"""
import numpy as np
import pytest

import oracle as O
from cases import random_draws
from mcm_amd import swimmer as SW

pytestmark = pytest.mark.gpu

H = 0.1


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _total_momentum(m, st=None):
    """Solvent momentum (diagnostics) plus the beads' sum of m v."""
    p = np.asarray(m.diagnostics()["momentum"], dtype=np.float64)
    if st is not None:
        p = p + (st["vel"][:, :3] * st["vel"][:, 3:4]).sum(axis=0)
    return p


def _shell_force_torque(st, nshell=643):
    x, f = st["pos"][:nshell], st["force"][:nshell]
    return f.sum(0), np.cross(x - x[0], f).sum(0)


def _compare_forms(st32, st64, tag, tol_pos=1e-6, tol_vel=1e-5):
    dp = np.abs(st32["pos"] - st64["pos"]).max()
    dv = np.abs(st32["vel"][:, :3] - st64["vel"][:, :3]).max()
    f32, t32 = _shell_force_torque(st32)
    f64, t64 = _shell_force_torque(st64)
    ef = np.abs(f32 - f64).max() / np.abs(f64).max()
    et = np.abs(t32 - t64).max() / np.abs(t64).max()
    print(f"[result] {tag}: bead position error {dp:.2e}, velocity error {dv:.2e}, shell force {ef:.2e} and "
          f"torque {et:.2e} relative", flush=True)
    assert dp < tol_pos and dv < tol_vel and ef < 1e-5 and et < 1e-5


def _vel_err(v32, v64):
    """Per-particle velocity error of the FP32-bulk form, norm-wise relative
    to the thermal scale: |dv| / max(|v|, 1).  The largest come from cells
    of two particles with a large thermostat factor (FP32 map) and from the
    particle that takes its cell's momentum residual."""
    d = np.linalg.norm(v32[:, :3] - v64[:, :3], axis=1)
    return d / np.maximum(np.linalg.norm(v64[:, :3], axis=1), 1.0)


def _representable_fluid(L, a, rng):
    """Uniform positions and unit normal velocities, stored exactly in both
    forms (encoded as FP32-bulk records at shift a and decoded): a created
    or uploaded FP64 state differs from its FP32-bulk copy by the FP32
    rounding of the offsets, which moves a particle within 2^-25 below a
    cell face into the next cell (about 15 of 168M), so forms compared
    cell by cell start from this one."""
    n = 10 * int(np.prod(L))
    pos = rng.random((n, 3), dtype=np.float32).astype(np.float64) * np.array(L, dtype=np.float64)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.standard_normal((n, 3), dtype=np.float32)
    cell, off, w = O.f32_encode(L, a, pos, vel)
    return O.f32_decode(L, a, cell, off, w)


# ---------------------------------------------------------------------- C0
C0_L = (32, 32, 32)


def _c0_swimmer():
    sw = SW.clami(levels=3, radius=5.0, center=(16.0, 16.0, 16.0), n_cilia=16, rodlength=10, seed=1)
    assert len(sw["pos3"]) == 1123
    return sw


def test_c0_swimmer_bit_exact_with_reference_draws(mcm):
    L, dt = C0_L, H / 50
    _print_case("C0_bit_exact", L=L, beads=1123, substeps=50)
    mt = O.MTRand(5489)
    a = np.array([mt.dblexc() for _ in range(3)])
    n = 10 * int(np.prod(L))
    pos, vel = O.fluid_create(mt, L, n)
    sw = _c0_swimmer()
    o = O.MDOracle(sw, globalt=0.0, ststep=dt, tstep=H)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True) as m:
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        m.md_attach(sw, globalt=0.0, ststep=dt)
        for s in range(3):
            cfg, keep, rec = O.make_cfg(L, rng=O.RNG_MT, mt=mt, record=True)
            O.collide(cfg, a, pos, vel, o.pos, o.vel)
            a_next = np.array([mt.dblexc() for _ in range(3)])
            O.stream(pos, vel, H)
            o.step()
            m.step(H, next_shift=a_next, inj_gamma=np.nan_to_num(rec[0], nan=1.0),
                   inj_axis=np.nan_to_num(rec[1], nan=0.0))
            st = m.md_fetch()
            gp, gv = m.download_fluid()
            assert m.diagnostics()["flags"] == 0
            assert np.array_equal(gp, pos), f"step {s}: solvent positions differ"
            assert np.array_equal(gv[:, :3], vel[:, :3]), f"step {s}: solvent velocities differ"
            assert np.array_equal(st["pos"], o.pos), f"step {s}: bead positions differ"
            assert np.array_equal(st["vel"][:, :3], o.vel[:, :3]), f"step {s}: bead velocities differ"
            assert st["globalt"] == o.globalt
            a = a_next


def test_c0_swimmer_1000_steps(mcm):
    L, dt, steps = C0_L, H / 50, 1000
    _print_case("C0_long", L=L, steps=steps)
    n = 10 * int(np.prod(L))
    sw = _c0_swimmer()
    with mcm.MPCD(L, precision=mcm.FP64, seed=1) as m:
        m.create_fluid(n)
        m.md_attach(sw, globalt=0.0, ststep=dt)
        m.step(H)  # the diagnostics hold the solvent's sums from the first step on
        p0 = _total_momentum(m, m.md_fetch())
        for _ in range(steps):
            m.step(H)
        st = m.md_fetch()
        p1 = _total_momentum(m, st)
        d = m.diagnostics()
    drift = np.abs(p1 - p0).max() / (n + len(sw["pos3"]))
    kT = d["twice_kinetic"] / (3.0 * n)
    print(f"[result] C0 {steps} steps: momentum drift per particle {drift:.2e}, kT {kT:.4f}, "
          f"globalt {st['globalt']:.3f}", flush=True)
    assert d["flags"] == 0 and d["n"] == n
    assert drift < 1e-12
    assert abs(kT - 1.0) < 0.01
    assert st["globalt"] == pytest.approx((steps + 1) * H, rel=1e-9)


# ---------------------------------------------------------------------- C1
C1_L = (256, 256, 256)
C1_N = 10 * 256 ** 3
CALPHA, SALPHA = -0.6427876, 0.76604444  # src/libcilia/mpcd.h:74-75


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_c1_invariants_20_steps(mcm, prec_name):
    _print_case("C1_invariants", prec=prec_name, n=C1_N)
    with mcm.MPCD(C1_L, precision=getattr(mcm, prec_name), seed=2) as m:
        m.create_fluid(C1_N)
        m.step(H)
        d0 = m.diagnostics()
        for _ in range(20):
            m.step(H)
        d = m.diagnostics()
    kT = d["twice_kinetic"] / (3.0 * C1_N)
    print(f"[result] C1 {prec_name}: kT {kT:.5f}, momentum {d['momentum']}", flush=True)
    assert d0["flags"] == 0 and d["flags"] == 0
    assert d["n"] == C1_N
    assert abs(kT - 1.0) < 2e-3
    if prec_name == "FP32BULK":
        assert np.array_equal(d["momentum_q"], d0["momentum_q"])
    else:
        assert np.abs(np.asarray(d["momentum"]) - np.asarray(d0["momentum"])).max() / C1_N < 1e-14


def test_c1_fp32_against_fp64_one_step(mcm):
    """Both forms take one step from the same state (one both store
    exactly) with the same shift and draws."""
    _print_case("C1_fp32_vs_fp64", n=C1_N)
    rng = np.random.default_rng(8)
    a, a_next = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    pos, vel = _representable_fluid(C1_L, a, rng)
    g, ax = random_draws(256 ** 3, rng)
    out = {}
    for name in ("FP32BULK", "FP64"):
        with mcm.MPCD(C1_L, precision=getattr(mcm, name), track_ids=True) as m:
            m.set_shift(a)
            m.upload_fluid(pos, vel)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            assert m.diagnostics()["flags"] == 0
            out[name] = m.download_fluid()
    del pos, vel
    p32, v32 = out["FP32BULK"]
    p64, v64 = out["FP64"]
    err = _vel_err(v32, v64)
    dx = p32 - p64
    dx -= 256.0 * np.round(dx / 256.0)
    print(f"[result] C1 one step FP32-bulk vs FP64: velocity error max {err.max():.2e}, 99% "
          f"{np.quantile(err, 0.99):.2e}; position {np.abs(dx).max():.2e}", flush=True)
    assert err.max() < 1e-5 and np.quantile(err, 0.99) < 1e-6
    assert np.abs(dx).max() < 1e-5


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_c1_energy_without_thermostat(mcm, prec_name):
    """Exact rotation constants keep the kinetic energy to rounding; the
    reference's truncated ones (cos^2 + sin^2 = 1 - 1.743e-8) lose
    (cos^2 + sin^2 - 1) x 2/3 x 0.9 of it per step (the share normal to the
    axis of the relative energy, 1 - cells/N)."""
    steps = 20
    a = np.deg2rad(130.0)
    res = {}
    for name, c, s_ in (("exact", np.cos(a), np.sin(a)), ("truncated", CALPHA, SALPHA)):
        _print_case("C1_energy", prec=prec_name, constants=name, steps=steps)
        with mcm.MPCD(C1_L, precision=getattr(mcm, prec_name), seed=12, thermostat=False, calpha=c,
                      salpha=s_) as m:
            m.create_fluid(C1_N)
            m.step(H)
            e0 = m.diagnostics()["twice_kinetic"]
            for _ in range(steps):
                m.step(H)
            d = m.diagnostics()
        assert d["flags"] == 0
        res[name] = (d["twice_kinetic"] - e0) / e0
    predicted = (1.0 + (CALPHA ** 2 + SALPHA ** 2 - 1.0) * (2.0 / 3.0) * 0.9) ** steps - 1.0
    print(f"[result] C1 {prec_name} energy change over {steps} steps without thermostat: exact "
          f"{res['exact']:.3e}, truncated {res['truncated']:.4e} (predicted {predicted:.4e})", flush=True)
    assert abs(res["exact"]) < (1e-13 if prec_name == "FP64" else 1e-8)
    assert abs(res["truncated"] / predicted - 1.0) < 0.02


# ---------------------------------------------------------------------- C2
C2_L = (100, 100, 100)


def _c2_swimmer():
    sw = SW.clami(levels=3, radius=16.0, center=(50.0, 50.0, 50.0), n_cilia=400, rodlength=20, seed=2)
    assert len(sw["pos3"]) == 24643
    return sw


def test_c2_fp32_against_fp64_one_step(mcm):
    L, dt = C2_L, H / 50
    _print_case("C2_fp32_vs_fp64", L=L, beads=24643)
    rng = np.random.default_rng(21)
    a, a_next = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    pos, vel = _representable_fluid(L, a, rng)
    g, ax = random_draws(int(np.prod(L)), rng)
    sw = _c2_swimmer()
    st, fl = {}, {}
    for name in ("FP32BULK", "FP64"):
        with mcm.MPCD(L, precision=getattr(mcm, name), track_ids=True) as m:
            m.set_shift(a)
            m.upload_fluid(pos, vel)
            m.md_attach(sw, globalt=0.0, ststep=dt)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax)
            assert m.diagnostics()["flags"] == 0
            st[name] = m.md_fetch()
            fl[name] = m.download_fluid()[1]
    _compare_forms(st["FP32BULK"], st["FP64"], "C2 one step")
    err = _vel_err(fl["FP32BULK"], fl["FP64"])
    print(f"[result] C2 one step: solvent velocity error max {err.max():.2e}, 99% {np.quantile(err, 0.99):.2e}",
          flush=True)
    assert err.max() < 1e-5 and np.quantile(err, 0.99) < 1e-6


def test_c2_swimming_velocity_fp32_against_fp64(mcm):
    L, dt, blocks, per = C2_L, H / 50, 10, 200
    n = 10 * int(np.prod(L))
    sw = _c2_swimmer()
    vel = {}
    for name in ("FP32BULK", "FP64"):
        _print_case("C2_swimming", prec=name, steps=blocks * per)
        with mcm.MPCD(L, precision=getattr(mcm, name), seed=31) as m:
            m.create_fluid(n)
            m.md_attach(sw, globalt=0.0, ststep=dt)
            x = [m.md_fetch()["pos"][:643].mean(0)]
            for _ in range(blocks):
                for _ in range(per):
                    m.step(H)
                x.append(m.md_fetch()["pos"][:643].mean(0))
            assert m.diagnostics()["flags"] == 0
        vel[name] = np.diff(np.array(x), axis=0) / (per * H)
    m32, m64 = vel["FP32BULK"].mean(0), vel["FP64"].mean(0)
    se = np.sqrt(vel["FP32BULK"].var(0, ddof=1) / blocks + vel["FP64"].var(0, ddof=1) / blocks)
    print(f"[result] C2 shell velocity over {blocks * per} steps: FP32-bulk {m32}, FP64 {m64}, "
          f"standard error {se}", flush=True)
    assert np.all(np.abs(m32 - m64) <= 5.0 * se + 1e-4)


# ---------------------------------------------------------------------- C3
def test_c3_ensemble_64_boxes(mcm):
    """The ensemble of the benchmark (seed 7: edges, cilia counts, rod
    lengths and radii drawn as there)."""
    steps, dt = 20, H / 50
    rng = np.random.default_rng(7)
    prec = mcm.FP32BULK
    variant, _ = mcm.autotune((48, 48, 48), 10 * 48 ** 3, steps=4, precision=prec, seed=7)
    specs = []
    for i in range(64):
        L = int(rng.integers(32, 65))
        ncil, rod, R = int(rng.integers(8, 129)), int(rng.integers(10, 21)), float(rng.uniform(5.0, 10.0))
        specs.append((L, ncil, rod, R, 1000 + i))
    _print_case("C3_ensemble", boxes=64, variant=variant, steps=steps)

    def make(spec):
        L, ncil, rod, R, seed = spec
        m = mcm.MPCD((L, L, L), precision=prec, seed=seed, tile_variant=variant, track_ids=True)
        m.create_fluid(10 * L ** 3)
        sw = SW.clami(levels=3, radius=R, center=(L / 2.0,) * 3, n_cilia=ncil, rodlength=rod, seed=seed)
        m.md_attach(sw, globalt=0.0, ststep=dt)
        return m

    edges = [s[0] for s in specs]
    sample = sorted({int(np.argmin(edges)), int(np.argmax(edges)), 17})
    boxes = []
    try:
        for s in specs:
            boxes.append(make(s))
        with mcm.Ensemble(boxes) as e:
            e.step(H, steps)
            e.sync()
            flags = [b.diagnostics()["flags"] for b in boxes]
            got = {i: (boxes[i].download_fluid(), boxes[i].md_fetch()) for i in sample}
    finally:
        for b in boxes:
            b.close()
    assert not any(flags), f"flags {flags}"
    for i in sample:
        with make(specs[i]) as m:
            for _ in range(steps):
                m.step(H)
            (p, v), st = m.download_fluid(), m.md_fetch()
        (gp, gv), gst = got[i]
        assert np.array_equal(gp, p) and np.array_equal(gv, v), f"box {i}: solvent differs"
        assert np.array_equal(gst["pos"], st["pos"]) and np.array_equal(gst["vel"], st["vel"]), f"box {i}: beads"


# ---------------------------------------------------------------------- C4
C4_L = (128, 128, 128)


def _c4_swimmer():
    sw = SW.clami(levels=3, radius=20.0, center=(64.0, 64.0, 64.0), n_cilia=1000, rodlength=20, seed=4)
    assert len(sw["pos3"]) == 60643
    return sw


def test_c4_fp32_against_fp64_one_step(mcm):
    L, dt = C4_L, H / 100
    _print_case("C4_fp32_vs_fp64", L=L, beads=60643, substeps=100)
    rng = np.random.default_rng(44)
    a, a_next = rng.uniform(0, 1, 3), rng.uniform(0, 1, 3)
    pos, vel = _representable_fluid(L, a, rng)
    sw = _c4_swimmer()
    ran = rng.normal(size=3 * (len(pos) + len(sw["pos3"])))
    st = {}
    for name in ("FP32BULK", "FP64"):
        with mcm.MPCD(L, precision=getattr(mcm, name), track_ids=True, collision=mcm.AT_A) as m:
            m.set_shift(a)
            m.upload_fluid(pos, vel)
            m.md_attach(sw, globalt=0.0, ststep=dt)
            m.step(H, next_shift=a_next, inj_ran=ran)
            assert m.diagnostics()["flags"] == 0
            st[name] = m.md_fetch()
    _compare_forms(st["FP32BULK"], st["FP64"], "C4 one step")


def _cells(L, a, pos):
    c = np.floor(pos + a).astype(np.int64) % np.asarray(L)
    return c[:, 0] + L[0] * (c[:, 1] + L[1] * c[:, 2])


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_c4_long_run_drift(mcm, prec_name):
    L, dt, steps = C4_L, H / 100, 2000
    _print_case("C4_long", prec=prec_name, steps=steps)
    n = 10 * int(np.prod(L))
    sw = _c4_swimmer()
    with mcm.MPCD(L, precision=getattr(mcm, prec_name), seed=41, collision=mcm.AT_A, track_ids=True) as m:
        m.create_fluid(n)
        m.md_attach(sw, globalt=0.0, ststep=dt)
        m.step(H)  # the diagnostics hold the solvent's sums from the first step on
        p0 = _total_momentum(m, m.md_fetch())
        for k in range(steps):
            m.step(H)
            if k % 500 == 499:
                print(f"[progress] C4 {prec_name} step {k + 1}", flush=True)
        st = m.md_fetch()
        p1 = _total_momentum(m, st)
        d = m.diagnostics()
        # the next collision's angular momentum, cell by cell, in the cells
        # the beads leave alone
        a = m.get_shift()
        pos, v0 = m.download_fluid()
        m.step(H)
        _, v1 = m.download_fluid()
        assert m.diagnostics()["flags"] == 0
    drift = np.abs(p1 - p0).max() / (n + len(sw["pos3"]))
    kT = d["twice_kinetic"] / (3.0 * n)
    cell = _cells(L, a, pos)
    bead_cells = np.unique(_cells(L, a, st["pos"]))
    free = ~np.isin(cell, bead_cells)
    # the collision's lever arm: the offset in the particle's cell rounded to
    # FP32, and in the FP32-bulk form further to the 2^-24 grid of its
    # exact integer sums
    r = ((pos + a) - np.floor(pos + a)).astype(np.float32).astype(np.float64)
    if prec_name == "FP32BULK":
        r = np.rint(r * 2.0 ** 24) / 2.0 ** 24
    dL = np.cross(r[free], v1[free, :3] - v0[free, :3]).sum(0)
    scale = np.abs(np.cross(r[free], v1[free, :3] - v0[free, :3])).sum()
    print(f"[result] C4 {prec_name} {steps} steps: momentum drift per particle {drift:.2e}, kT {kT:.4f}; "
          f"one collision: angular momentum change {np.abs(dL).max():.2e} of {scale:.2e} exchanged "
          f"({free.sum()} particles in bead-free cells)", flush=True)
    assert d["flags"] == 0 and d["n"] == n
    assert drift < 1e-12
    assert abs(kT - 1.0) < 0.01
    assert np.abs(dL).max() < (1e-13 if prec_name == "FP64" else 1e-9) * scale
