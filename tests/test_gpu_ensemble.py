"""GPU parity of ensembles: independent boxes stepped together.

Every box of an ensemble (its own size, seed, fluid and optional swimmer)
must end bit-identical to the same box stepped alone with mpcd_step: the
solvent (positions and velocities by particle id) and the swimmer
(positions, velocities, rest lengths, beat state, globalt); the
diagnostics to 1e-12.  This holds for graph replays and for plain
launches, for both MD paths (a launch per phase over all swimmers, or a
workgroup per swimmer), across a change of the step length, and when the
boxes continue alone after the ensemble is released.
"""
import numpy as np
import pytest

from mcm_amd import swimmer as SW

pytestmark = pytest.mark.gpu

H = 0.1
DT = 0.002

# (size, seed, swimmer: (radius, n_cilia, rodlength) or None)
BOXES = [
    ((20, 16, 24), 3, (4.0, 6, 4)),
    ((16, 16, 16), 4, None),
    ((24, 20, 16), 5, (3.5, 10, 5)),
    ((13, 17, 11), 6, None),
]


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _make(mcm, prec, size, seed, sw, collision=None):
    kw = {} if collision is None else {"collision": collision}
    m = mcm.MPCD(size, precision=prec, track_ids=True, seed=seed, **kw)
    m.create_fluid(10 * int(np.prod(size)))
    if sw is not None:
        radius, n_cilia, rodlength = sw
        c = np.asarray(size, dtype=np.float64) / 2.0
        desc = SW.clami(levels=3, radius=radius, center=tuple(c), n_cilia=n_cilia, rodlength=rodlength, seed=seed)
        m.md_attach(desc, globalt=0.0, ststep=DT)
    return m


def _state(m, has_sw):
    pos, vel = m.download_fluid()
    d = m.diagnostics()
    st = {"pos": pos, "vel": vel, "mom": np.asarray(d["momentum"]), "e2": d["twice_kinetic"], "flags": d["flags"]}
    if has_sw:
        st["md"] = m.md_fetch()
    return st


def _assert_same(a, b, tag):
    assert a["flags"] == 0 and b["flags"] == 0
    assert np.array_equal(a["pos"], b["pos"]), f"{tag}: solvent positions differ, max {np.abs(a['pos'] - b['pos']).max()}"
    assert np.array_equal(a["vel"], b["vel"]), f"{tag}: solvent velocities differ"
    # the diagnostics' per-tile sums are atomics (order varies): 1e-12
    np.testing.assert_allclose(a["mom"], b["mom"], rtol=0, atol=1e-9)
    assert abs(a["e2"] - b["e2"]) <= 1e-12 * abs(b["e2"]), f"{tag}: diagnostics differ"
    if "md" in a:
        for k in ("pos", "vel", "newlength", "oldlength", "beat"):
            assert np.array_equal(a["md"][k], b["md"][k]), f"{tag}: swimmer {k} differs"
        assert a["md"]["globalt"] == b["md"]["globalt"]


@pytest.mark.parametrize("prec_name,graph,md_wide", [("FP32BULK", True, True), ("FP64", True, True),
                                                    ("FP32BULK", False, False), ("FP64", False, True),
                                                    ("FP64", True, False)])
def test_ensemble_bit_identical_to_boxes_alone(mcm, prec_name, graph, md_wide):
    prec = getattr(mcm, prec_name)
    _print_case("ensemble", prec=prec_name, graph=graph, md_wide=md_wide, boxes=len(BOXES))
    alone = [_make(mcm, prec, *b) for b in BOXES]
    ens_boxes = [_make(mcm, prec, *b) for b in BOXES]
    try:
        for m in alone:
            for _ in range(4):
                m.step(H)
            m.step(0.5 * H)
        with mcm.Ensemble(ens_boxes, graph=graph, md_wide=md_wide) as e:
            e.step(H, 3)
            e.step(H, 1)
            e.step(0.5 * H, 1)  # another step length: tables and graph rebuilt
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i}")
        # released: each box continues alone from where the ensemble left it
        for m, mb in zip(alone, ens_boxes):
            m.step(H)
            mb.step(H)
        for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
            _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i} after")
    finally:
        for m in alone + ens_boxes:
            m.close()


def test_ensemble_refresh_schedules_that_change_every_step(mcm):
    """Steps of 0.025 in sub-steps of 0.002: the swimmers' rest-length
    refresh schedules cycle through five patterns, each replayed from a
    graph of its own; every box bit-identical to the box stepped alone."""
    prec = mcm.FP64
    tstep, steps = 0.025, 12
    _print_case("ensemble_schedules", tstep=tstep, steps=steps)
    alone = [_make(mcm, prec, *b) for b in BOXES]
    ens_boxes = [_make(mcm, prec, *b) for b in BOXES]
    try:
        for m in alone:
            for _ in range(steps):
                m.step(tstep)
        with mcm.Ensemble(ens_boxes) as e:
            e.step(tstep, steps)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i}")
    finally:
        for m in alone + ens_boxes:
            m.close()


@pytest.mark.parametrize("prec_name,rule", [("FP32BULK", "AT_A"), ("FP64", "AT_A"), ("FP32BULK", "AT")])
def test_ensemble_mpc_at_bit_identical_to_boxes_alone(mcm, prec_name, rule):
    """MPC-AT (+a) boxes, with and without swimmers, stepped as an ensemble
    end bit-identical to the same boxes stepped alone."""
    prec, coll = getattr(mcm, prec_name), getattr(mcm, rule)
    _print_case("ensemble_at", prec=prec_name, rule=rule, boxes=len(BOXES))
    alone = [_make(mcm, prec, *b, collision=coll) for b in BOXES]
    ens_boxes = [_make(mcm, prec, *b, collision=coll) for b in BOXES]
    try:
        for m in alone:
            for _ in range(3):
                m.step(H)
        with mcm.Ensemble(ens_boxes) as e:
            e.step(H, 3)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i}")
    finally:
        for m in alone + ens_boxes:
            m.close()


@pytest.mark.parametrize("md_block", [256, 1024])
def test_ensemble_md_workgroup_sizes_bit_identical(mcm, md_block):
    """The wide MD launches at another workgroup size give the same bits."""
    prec = mcm.FP32BULK
    _print_case("ensemble_md_block", md_block=md_block)
    alone = [_make(mcm, prec, *b) for b in BOXES]
    ens_boxes = [_make(mcm, prec, *b) for b in BOXES]
    try:
        for m in alone:
            for _ in range(3):
                m.step(H)
        with mcm.Ensemble(ens_boxes) as e:
            e.set_md_block(md_block)
            e.step(H, 3)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i}")
            with pytest.raises(mcm.MPCDError):
                e.set_md_block(300)
    finally:
        for m in alone + ens_boxes:
            m.close()


def test_ensemble_autotune_keeps_the_run(mcm):
    """The autotuner's timed steps are steps of the run: after it, the boxes
    equal the boxes stepped alone as often, and it reports every candidate."""
    prec = mcm.FP32BULK
    _print_case("ensemble_autotune")
    alone = [_make(mcm, prec, *b) for b in BOXES]
    ens_boxes = [_make(mcm, prec, *b) for b in BOXES]
    try:
        steps = 1
        taken = 4 * (steps + 1)
        for m in alone:
            for _ in range(taken + 2):
                m.step(H)
        with mcm.Ensemble(ens_boxes) as e:
            r = e.autotune(H, steps=steps)
            assert r["md_block"] in (256, 512, 1024) and r["ms_per_step"] > 0
            assert min(r["candidates_ms"].values()) == pytest.approx(r["ms_per_step"])
            e.step(H, 2)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                _assert_same(_state(mb, BOXES[i][2] is not None), _state(m, BOXES[i][2] is not None), f"box {i}")
    finally:
        for m in alone + ens_boxes:
            m.close()


SPHERE_BOXES = [((24, 24, 24), 11, 4.0), ((20, 22, 18), 12, 3.0), ((26, 20, 24), 13, 3.5)]


def _sphere_box(mcm, prec, size, seed, radius):
    """A box with a rigid sphere at its centre and fluid outside it."""
    rng = np.random.default_rng(seed)
    c = np.asarray(size, dtype=np.float64) / 2.0
    n = 10 * int(np.prod(size))
    pos = rng.random((n, 3)) * np.asarray(size, dtype=np.float64)
    pos = pos[((pos - c) ** 2).sum(1) > radius * radius]
    vel = np.ones((len(pos), 4))
    vel[:, :3] = rng.normal(size=(len(pos), 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    m = mcm.MPCD(size, precision=prec, track_ids=True, seed=seed)
    m.upload_fluid(pos, vel)
    m.sphere_attach(radius, 200.0, tuple(c), v=(0.05, 0.0, 0.02), w=(0.0, 0.01, 0.0))
    return m


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_ensemble_rigid_spheres_match_boxes_alone(mcm, prec_name):
    """Boxes with a rigid sphere each, stepped as an ensemble: the first step
    bit-identical to the boxes stepped alone (the sphere's state to 1e-12:
    its impulse is a sum over tiles in atomic order), the next steps equal
    to roundoff carried by that sphere state (FP64 1e-10; FP32-bulk: a
    velocity may land one 2^-21 grid step apart)."""
    prec = getattr(mcm, prec_name)
    _print_case("ensemble_spheres", prec=prec_name, boxes=len(SPHERE_BOXES))
    alone = [_sphere_box(mcm, prec, *b) for b in SPHERE_BOXES]
    ens_boxes = [_sphere_box(mcm, prec, *b) for b in SPHERE_BOXES]
    tol = 1e-10 if prec_name == "FP64" else 2e-6
    try:
        for m in alone:
            m.step(H)
        with mcm.Ensemble(ens_boxes) as e:
            e.step(H, 1)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                a, b = _state(mb, False), _state(m, False)
                assert a["flags"] == 0 and b["flags"] == 0
                assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"]), f"box {i}"
                sa, sb = mb.sphere_state(), m.sphere_state()
                for k in ("x", "v", "w", "force", "torque"):
                    np.testing.assert_allclose(sa[k], sb[k], rtol=1e-12, atol=1e-15, err_msg=f"box {i} sphere {k}")
            for m in alone:
                for _ in range(3):
                    m.step(H)
            e.step(H, 3)
            e.sync()
            for i, (m, mb) in enumerate(zip(alone, ens_boxes)):
                a, b = _state(mb, False), _state(m, False)
                assert a["flags"] == 0 and b["flags"] == 0
                assert np.abs(a["pos"] - b["pos"]).max() < tol and np.abs(a["vel"] - b["vel"]).max() < tol, f"box {i}"
                sa, sb = mb.sphere_state(), m.sphere_state()
                np.testing.assert_allclose(sa["v"], sb["v"], rtol=1e-9, atol=1e-12)
                assert abs(sa["v"][0]) > 0.0
    finally:
        for m in alone + ens_boxes:
            m.close()


def test_ensemble_refuses_mixed_rules_and_fp32_walls(mcm):
    _print_case("ensemble_rule_refusals")
    a = _make(mcm, mcm.FP32BULK, (16, 16, 16), 1, None, collision=mcm.AT_A)
    b = _make(mcm, mcm.FP32BULK, (16, 16, 16), 2, None, collision=mcm.AT)
    w = mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK, seed=3, walls=mcm.NOSLIP)
    w.create_fluid(10 * 16 ** 3)
    try:
        with pytest.raises(mcm.MPCDError):
            mcm.Ensemble([a, b])  # two collision rules
        with pytest.raises(mcm.MPCDError):
            mcm.Ensemble([w])  # FP32-bulk walls are not in the ensemble kernels
        s1 = _sphere_box(mcm, mcm.FP32BULK, (16, 16, 16), 4, 3.0)
        s2 = _make(mcm, mcm.FP32BULK, (16, 16, 16), 5, None)
        try:
            with pytest.raises(mcm.MPCDError):
                mcm.Ensemble([s1, s2])  # a sphere in some boxes only
        finally:
            s1.close()
            s2.close()
    finally:
        for m in (a, b, w):
            m.close()


def test_ensemble_members_refuse_state_changes(mcm):
    _print_case("ensemble_refusals")
    a = _make(mcm, mcm.FP32BULK, (16, 16, 16), 1, None)
    b = _make(mcm, mcm.FP32BULK, (16, 16, 16), 2, None)
    c = _make(mcm, mcm.FP64, (16, 16, 16), 3, None)
    try:
        with pytest.raises(mcm.MPCDError):
            mcm.Ensemble([a, c])  # mixed precision
        with pytest.raises(mcm.MPCDError):
            mcm.Ensemble([a, a])  # a box twice
        e = mcm.Ensemble([a, b])
        with pytest.raises(mcm.MPCDError):
            a.step(H)
        with pytest.raises(mcm.MPCDError):
            a.create_fluid(100)
        with pytest.raises(mcm.MPCDError):
            mcm.Ensemble([b])  # already in one
        e.step(H, 2)
        e.sync()
        assert a.diagnostics()["flags"] == 0
        e.close()
        a.step(H)  # usable alone again
    finally:
        for m in (a, b, c):
            m.close()


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_create_fluid_reproducible(mcm, prec_name):
    """A seed gives the same fluid bits every time (the mean-velocity
    removal sums in a fixed order)."""
    prec = getattr(mcm, prec_name)
    _print_case("create_fluid_reproducible", prec=prec_name)
    a = _make(mcm, prec, (40, 36, 32), 9, None)
    b = _make(mcm, prec, (40, 36, 32), 9, None)
    try:
        pa, va = a.download_fluid()
        pb, vb = b.download_fluid()
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)
        # FP32-bulk stores velocities on the 2^-21 grid after the mean is removed
        assert np.abs(va[:, :3].mean(0)).max() < (1e-12 if prec_name == "FP64" else 1e-8)
    finally:
        a.close()
        b.close()
