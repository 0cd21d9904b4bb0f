"""Swimming of a ciliated swimmer: FP32-bulk against FP64 at fixed error.

A Clami built by the configuration driver's factory (radius 6, rings at
-45, 0, 45 degrees with 6, 8, 6 cilia of 16 beads, all beating in phase)
in a 48^3 box at the reference's kT = 1.  Its beads' centre of mass moves
forward in each power stroke (75 time units) and back in the recovery
stroke (150), netting a forward step per beat of 225.  Thermal noise moves
it too (Stokes diffusion, D ~ 1e-3), so the two forms, which draw
different random numbers, are compared statistically: over beats 2-5 (the
first is the start-up) the mean displacement of the power strokes, its
mean over whole beats and the swimming direction must agree within their
scatter.
"""

import numpy as np
import pytest

from mcm_amd import sim as S

pytestmark = pytest.mark.gpu

TSTEP = 0.05
PERIOD = 225.0
EVERY = 50  # steps between samples: 2.5 time units
CFG = """name = "swim"; globalt = 0.0; desiredt = 100000.0; omega = 0.0;
sizex = 48; sizey = 48; sizez = 48; runtime = 3600.0; tstep = 0.05;
md_objects = ( { name = "swim"; type = "Clami"; radius = 6.0; springconstant = 20000.0;
  position = [ 24.0, 24.0, 24.0 ]; y_thetas = [ -45.0, 0.0, 45.0 ]; num_equators = [ 6, 8, 6 ];
  theta_offset = [ 0.39269908169872414, 0.0, 0.39269908169872414 ];
  cilia = { type = "ConstantBeatCilia"; name = "c"; rodlength = 16; beatstrength = 0.17; bending = 0.01;
    min_bend = -0.7; max_bend = 0.7; m_speed = 5.3; springconstant = 20000.0; position = [0.0, 0.0, 0.0];
    thetaX = 0.0; thetaY = 0.0; thetaZ = 0.0; write_data = false; power_stroke_duration = 75.0;
    recovery_stroke_duration = 150.0; start_delay = 0.0; }; } );
"""


def _track(mcm, model, prec, seed, beats):
    """Bead centre of mass every EVERY steps over `beats` beats."""
    nb = model["pos3"].shape[0]
    steps_per_beat = int(round(PERIOD / TSTEP))
    xs = []
    with mcm.MPCD((48, 48, 48), precision=prec, seed=seed) as m:
        m.create_fluid(10 * 48 ** 3)
        m.md_attach(model, globalt=0.0, ststep=model["ststep"])
        for s in range(beats * steps_per_beat // EVERY + 1):
            if s:
                for _ in range(EVERY):
                    m.step(TSTEP)
            xs.append(m.md_fetch()["pos"].mean(axis=0))
        assert m.diagnostics()["flags"] == 0
    assert nb == len(model["pos3"])
    return np.array(xs)


def test_swimming_stroke_fp32_vs_fp64(mcm, tmp_path):
    path = tmp_path / "sim.cfg"
    path.write_text(CFG)
    with S.Simulation(str(path)) as s:
        model = s.model()
    beats = 5
    per_beat = int(round(PERIOD / TSTEP)) // EVERY  # samples per beat
    power = int(round(75.0 / TSTEP)) // EVERY
    res = {}
    for name, prec, seed in (("FP64", mcm.FP64, 41), ("FP32BULK", mcm.FP32BULK, 42)):
        print(f"[case] swimming prec={name} beats={beats}", flush=True)
        x = _track(mcm, model, prec, seed, beats)
        starts = [b * per_beat for b in range(1, beats)]
        stroke = np.array([x[i + power] - x[i] for i in starts])  # power stroke displacements
        net = np.array([x[i + per_beat] - x[i] for i in starts])  # whole beats
        res[name] = (stroke, net)
        print(f"[result] {name} power stroke {stroke.mean(0)} +- {stroke.std(0, ddof=1) / 2}, "
              f"per beat {net.mean(0)} +- {net.std(0, ddof=1) / 2}", flush=True)
    (s64, n64), (s32, n32) = res["FP64"], res["FP32BULK"]
    d64 = s64.mean(0) / np.linalg.norm(s64.mean(0))
    d32 = s32.mean(0) / np.linalg.norm(s32.mean(0))
    assert d64 @ d32 > 0.95  # the same stroke direction
    a64, a32 = s64 @ d64, s32 @ d64  # stroke length along it, per beat
    # standard error of the difference from the beats' scatter, with a floor of 0.2 per beat for the
    # scatter four beats can underestimate
    n = len(a64)
    sig = np.sqrt((max(a64.var(ddof=1), 0.2 ** 2) + max(a32.var(ddof=1), 0.2 ** 2)) / n)
    print(f"[result] stroke length fp64 {a64.mean():.3f} fp32 {a32.mean():.3f} (sigma {sig:.3f})", flush=True)
    assert a64.mean() > 1.0 and a32.mean() > 1.0
    assert abs(a64.mean() - a32.mean()) < 4.0 * sig
    b64, b32 = n64 @ d64, n32 @ d64
    sig_b = np.sqrt((max(b64.var(ddof=1), 0.3 ** 2) + max(b32.var(ddof=1), 0.3 ** 2)) / n)
    print(f"[result] net step per beat fp64 {b64.mean():.3f} fp32 {b32.mean():.3f} (sigma {sig_b:.3f})", flush=True)
    assert abs(b64.mean() - b32.mean()) < 4.0 * sig_b
