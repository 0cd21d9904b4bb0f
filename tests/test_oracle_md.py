"""Pins of the oracle's MD half (bead-spring sub-steps, forces, beat).

* Forces are -grad U: central finite differences of U = 1/2 sum over the
  one-sided spring terms of k (d - b)^2, with the terms enumerated here
  independently from the reference's loops (src/sphere.cpp:87-119,
  src/cilia.cpp:127-239, src/clami.cpp:188-198), match the oracle's force
  on every bead.  A dropped factor 2, a sign or a wrong partner fails it.
* Newton's third law: every spring comes in a pair, so the summed
  acceleration is zero and the bead momentum (forces are accelerations, so
  sum v) is conserved to roundoff over many steps.
* Velocity Verlet is time-reversible: with the beat frozen, n steps
  forward, velocities negated, n steps forward again returns to the start.
* Energy: with the beat frozen, the error of 1/2 sum |v|^2 + U shrinks
  about four-fold when the sub-step is halved (velocity Verlet is second
  order).
* The constant-beat state machine (src/ConstantBeatCilia.cpp:8-23) follows
  a trace worked out by hand.
"""
import numpy as np
import pytest

import md_cases as M
import oracle as O


def spring_terms(sw, newlength):
    """(i, j, b, k) one-sided terms of the swimmer, from the reference's loops."""
    t = []
    for s in sw["spheres"]:
        o, n = s["first"], s["npart"]
        for i in range(1, n):
            t.append((o, o + i, s["radius"], 0.1 * s["springconstant"]))
            t.append((o + i, o, s["radius"], 0.1 * s["springconstant"]))
        for i in range(1, n):
            for j in s["nei"][s["nei_start"][i]:s["nei_start"][i + 1]]:
                t.append((o + i, o + int(j), s["headbond"], s["springconstant"]))
    for c, nl in zip(sw["cilia"], newlength):
        R, o, bl, k = c["rodlength"], c["first"], c["bondlength"], c["springconstant"]
        cross = np.sqrt(2.0) * bl
        for i in range(R - 1):
            t += [(o + i, o + i + 1, nl[i], k), (o + i + 1, o + i, nl[i], k)]
            for r in (1, 2):
                t += [(o + r * R + i, o + r * R + i + 1, bl, k), (o + r * R + i + 1, o + r * R + i, bl, k)]
        for i in range(R):
            for a, b in ((0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)):
                t.append((o + i + a * R, o + i + b * R, bl, k))
        for i in range(R - 1):
            li = np.sqrt(bl * bl + bl * nl[i])
            t += [(o + i, o + i + R + 1, li, k), (o + i, o + i + 2 * R + 1, li, k), (o + i + R, o + i + 1, li, k),
                  (o + i + 2 * R, o + i + 1, li, k), (o + i + R, o + i + 2 * R + 1, cross, k),
                  (o + i + 2 * R, o + i + R + 1, cross, k)]
        for i in range(1, R):
            li = np.sqrt(bl * bl + bl * nl[i - 1])
            t += [(o + i, o + i + R - 1, li, k), (o + i, o + i + 2 * R - 1, li, k), (o + i + R, o + i - 1, li, k),
                  (o + i + 2 * R, o + i - 1, li, k), (o + i + R, o + i + 2 * R - 1, cross, k),
                  (o + i + 2 * R, o + i + R - 1, cross, k)]
    for b0, b1, l0 in sw["attach"]:
        t += [(b0, b1, l0, sw["att_k"]), (b1, b0, l0, sw["att_k"])]
    return t


def potential(x, terms):
    i, j, b, k = (np.array(z) for z in zip(*terms))
    d = np.linalg.norm(x[i.astype(int)] - x[j.astype(int)], axis=1)
    return 0.5 * np.sum(k * (d - b) ** 2)


def frozen(**kw):
    """A swimmer whose rest lengths never change: no beat, no stall."""
    sw = M.swimmer(beatstrength=0.0, **kw)
    for c in sw["cilia"]:
        c["bending"] = 1e9
    return sw


def test_forces_are_minus_grad_u():
    sw = M.swimmer(levels=1, n_cilia=3, rodlength=6, seed=3)
    rng = np.random.default_rng(1)
    sw["pos3"] = sw["pos3"] + rng.normal(0, 0.02, sw["pos3"].shape)  # off the rest shape
    o = O.MDOracle(sw, globalt=0.505)  # fmod(0.505 + 1e-4, 0.01) >= ststep: no rest-length refresh
    o.recalc()
    terms = spring_terms(sw, [c["newlength"] for c in sw["cilia"]])
    x0 = sw["pos3"].copy()
    h = 1e-6
    g = np.zeros_like(x0)
    for b in range(len(x0)):
        for a in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[b, a] += h
            xm[b, a] -= h
            g[b, a] = (potential(xp, terms) - potential(xm, terms)) / (2 * h)
    scale = np.abs(g).max()
    assert np.abs(o.force + g).max() < 1e-6 * scale
    # the total acceleration vanishes (every spring has its partner)
    assert np.abs(o.force.sum(axis=0)).max() < 1e-9 * scale


def test_momentum_conserved_with_beat():
    sw = M.swimmer(seed=4)
    o = O.MDOracle(sw)
    p0 = sw["vel4"][:, :3].sum(axis=0)
    for _ in range(10):
        o.step()
    assert np.abs(o.vel[:, :3].sum(axis=0) - p0).max() < 1e-9
    assert o.globalt == pytest.approx(0.5)


def test_time_reversible():
    sw = frozen(seed=5)
    o = O.MDOracle(sw)
    for _ in range(4):
        o.step()
    o.vel[:, :3] *= -1.0
    for _ in range(4):
        o.step()
    assert np.abs(o.pos - sw["pos3"]).max() < 1e-8
    assert np.abs(o.vel[:, :3] + sw["vel4"][:, :3]).max() < 1e-6


def test_energy_error_is_second_order():
    sw = frozen(seed=6)
    sw["pos3"] = sw["pos3"] + np.random.default_rng(2).normal(0, 0.01, sw["pos3"].shape)
    terms = spring_terms(sw, [c["newlength"] for c in sw["cilia"]])
    dev = []
    for dt in (0.001, 0.0005):
        o = O.MDOracle(sw, ststep=dt)
        e0 = 0.5 * np.sum(o.vel[:, :3] ** 2) + potential(o.pos, terms)
        worst = 0.0
        for _ in range(4):
            o.step()
            worst = max(worst, abs(0.5 * np.sum(o.vel[:, :3] ** 2) + potential(o.pos, terms) - e0) / e0)
        dev.append(worst)
    assert dev[0] < 0.05
    assert 2.5 < dev[0] / dev[1] < 6.0  # velocity Verlet: energy error ~ dt^2


def test_constant_beat_trace():
    """power 0.15, recovery 0.3, Tstroke 0.02, tstep 0.05, m_speed 5.3:
    the power stroke ends at the first call with globalt - Tstroke >= 0.15
    (midpnt 2 -> 3, then + tstep/m_speed on every call), and the recovery
    ends when globalt - Tstroke - 0.15 >= 0.3 (midpnt back to 2, Tstroke =
    globalt)."""
    sw = M.swimmer(n_cilia=1, seed=7)
    c = sw["cilia"][0]
    c.update(Tstroke=0.02, power_stroke_duration=0.15, recovery_stroke_duration=0.3, m_speed=5.3, midpnt=2.0)
    o = O.MDOracle(sw)
    trace = []
    for _ in range(12):
        o.step()
        trace.append((round(o.globalt, 12),) + o.cilia_state()[0])
    expect, mid, ts, g = [], 2.0, 0.02, 0.0
    for _ in range(12):
        g += 0.05
        if g - ts >= 0.15 - 1e-12 and mid == 2.0:
            mid += 1
        if g - ts - 0.15 >= 0.3 - 1e-12 and mid > 2.0:
            ts, mid = g, 2.0
        if mid > 2.0:
            mid += 0.05 / 5.3
        expect.append((round(g, 12), mid, ts))
    assert [t[0] for t in trace] == [e[0] for e in expect]
    assert np.allclose([t[1] for t in trace], [e[1] for e in expect], rtol=0, atol=1e-12)
    assert np.allclose([t[2] for t in trace], [e[2] for e in expect], rtol=0, atol=1e-12)
    assert any(t[1] > 2.0 for t in trace) and trace[-1][1] == 2.0  # a full power-recovery cycle
