"""Pins of the oracle's beat machinery, by values worked out by hand and by
properties of the reference's formulas, not by retyping them:

* Cilia::length (src/cilia.cpp:79-121): the power-stroke profile is
  bondlength - 2 beatstrength at the mid point, falls off as a Lorentzian
  1/(d^2 + 1) and is symmetric about the mid point; the recovery branch is
  bondlength + beatstrength/2 (1 - 1/(midpnt - num_anker - 2)) at its first
  bead and goes as t^2.5 in between (t = 1/2 gives 2^-2.5); the stall rule
  returns the profile, the actual length moved by `bending` towards it, or
  the old length, by the sign rule, on a table of hand-made cases.
* The refresh schedule (src/cilia.cpp:206-207): 100 refreshes per time
  unit, on every tenth (ststep 1e-3) or fifth (ststep 2e-3) sub-step, also
  far from globalt = 0.
* Cilia::localbend (src/cilia.cpp:68-78): zero for a straight filament,
  (R - 1) 2 sin(theta/2) (2 r0 - r1 - r2) for rods bent on concentric arcs,
  and the switch it triggers in Cilia::switch_logic (src/cilia.cpp:241-258).
A planted mistake in any of these (a dropped factor, a sign, an off-by-one
index, a missing offset in the fmod test) fails one of the tests below.
"""
import math

import numpy as np
import pytest

import oracle as O

BL = 0.5


def one_cilium(rodlength=12, midpnt=6.0, beatstrength=0.17, bending=1e9, num_anker=2, rod0_spacing=BL,
               oldlength=None, constant_beat=False, pos=None):
    """One cilium, no shell: rod 0 along z with the given bead spacing, rods
    1 and 2 beside it at the bond length."""
    R = rodlength
    if pos is None:
        pos = np.zeros((3 * R, 3))
        for k in range(R):
            pos[k] = (0.0, 0.0, k * rod0_spacing)
            pos[R + k] = (BL, 0.0, k * BL)
            pos[2 * R + k] = (0.5 * BL, BL * math.sqrt(3) / 2, k * BL)
    vel = np.zeros((3 * R, 4))
    vel[:, 3] = 5.0
    old = np.full(R, BL) if oldlength is None else np.asarray(oldlength, dtype=np.float64)
    c = dict(first=0, rodlength=R, num_anker=num_anker, constant_beat=constant_beat, bondlength=BL,
             beatstrength=beatstrength, bending=bending, min_bend=-0.7, max_bend=0.7, m_speed=5.3,
             springconstant=20000.0, power_stroke_duration=75.0, recovery_stroke_duration=150.0, midpnt=midpnt,
             Tstroke=0.0, newlength=np.full(R, BL), oldlength=old)
    return dict(pos3=pos, vel4=vel, mobile=None, spheres=[], cilia=[c], attach=[], att_k=20000.0)


def length(sw, i, **kw):
    o = O.MDOracle(sw, **kw)
    return o.cilium_length(0, i)


# ---------------------------------------------------------------- length(i)
def test_power_stroke_mid_point_and_lorentzian():
    beta, m = 0.17, 6.0
    sw = one_cilium(midpnt=m, beatstrength=beta)
    assert length(sw, int(m)) == pytest.approx(0.5 - 0.34, abs=1e-15)
    # symmetric about the mid point, where both sides use the power branch
    for j in (1, 2):
        assert length(sw, int(m) + j) == pytest.approx(length(sw, int(m) - j), abs=1e-15)
    # (bl - l)(d^2 + 1) = 2 beta at every distance d
    for d in range(0, 5):
        assert (BL - length(sw, int(m) + d)) * (d * d + 1) == pytest.approx(0.34, rel=1e-13)


def test_power_stroke_fractional_mid_point():
    # midpnt = 6.25: bead 5 (d = -1.25) and bead 7 (d = 0.75)
    sw = one_cilium(midpnt=6.25, beatstrength=0.2)
    assert length(sw, 5) == pytest.approx(0.5 - 0.4 / (1.5625 + 1.0), abs=1e-15)
    assert length(sw, 7) == pytest.approx(0.5 - 0.4 / (0.5625 + 1.0), abs=1e-15)
    # bead 4 is at midpnt - 2.25 < midpnt - 2: the recovery branch, above bl
    assert length(sw, 4) > BL


def test_recovery_branch_anchor_and_exponent():
    beta, na = 0.2, 2
    # first bead of the recovery branch, i = num_anker + 1: t = 1
    sw = one_cilium(rodlength=13, midpnt=10.0, beatstrength=beta, num_anker=na)
    assert length(sw, na + 1) == pytest.approx(0.5 + beta * 5.0 / 12.0, abs=1e-15)  # 1 - 1/(10 - 2 - 2) = 5/6
    # far mid point (the smoothing factor is 1 - 1e-12): R - na - 1 = 10, so
    # i = na + 6 is t = 1/2 and the profile is 2^-2.5 of the anchor's
    sw = one_cilium(rodlength=13, midpnt=1e12, beatstrength=beta, num_anker=na)
    a = length(sw, na + 1) - BL
    assert a == pytest.approx(beta / 2, rel=1e-11)
    assert (length(sw, na + 6) - BL) / a == pytest.approx(0.17677669529663687, rel=1e-11)
    # t = 0.3 and t = 0.9: the exponent from the log ratio
    r = (length(sw, na + 8) - BL) / (length(sw, na + 2) - BL)
    assert math.log(r) / math.log(0.3 / 0.9) == pytest.approx(2.5, rel=1e-10)


# Stall rule, l = 0.16 (power stroke at the mid point, beta 0.17), bending
# 0.01; the actual length tl from the rod's bead spacing, the old length set.
STALL_CASES = [
    # tl,   old,  expected: which candidate
    (0.165, 0.5, 0.16),   # |l - tl| <= bending: the profile
    (0.155, 0.0, 0.16),
    (0.5, 0.5, 0.49),     # l below tl: tl - bending, towards l from old
    (0.5, 0.3, 0.3),      # tl - bending beyond old away from l: old
    (0.1, 0.05, 0.11),    # l above tl: tl + bending
    (0.1, 0.13, 0.13),    # tl + bending on the other side of old: old
]


@pytest.mark.parametrize("tl,old,expect", STALL_CASES)
def test_stall_rule_table(tl, old, expect):
    m = 6
    sw = one_cilium(midpnt=float(m), beatstrength=0.17, bending=0.01, rod0_spacing=tl,
                    oldlength=np.full(12, old))
    assert length(sw, m) == pytest.approx(expect, abs=1e-12)


# ----------------------------------------------------------------- refresh
@pytest.mark.parametrize("ststep,every,g0", [(1e-3, 10, 0.0), (2e-3, 5, 0.0), (1e-3, 10, 225.0)])
def test_refresh_schedule(ststep, every, g0):
    # globalt advanced by repeated additions, as the step loop does
    g, hits = g0, []
    n = int(round(1.0 / ststep))
    for k in range(n):
        if O.refresh_due(g, ststep):
            hits.append(k)
        g += ststep
    assert len(hits) == 100
    assert hits == list(range(0, n, every))


# --------------------------------------------------------------- localbend
def bent_cilium(R=10, theta=0.1, r=(11.0, 10.5, 10.5)):
    pos = np.zeros((3 * R, 3))
    for j in range(3):
        for i in range(R):
            pos[j * R + i] = (r[j] * math.cos(i * theta), r[j] * math.sin(i * theta), (0.0, 0.25, -0.25)[j])
    return one_cilium(rodlength=R, pos=pos)


def test_localbend_straight_is_zero():
    o = O.MDOracle(one_cilium())
    assert abs(o.localbend(0)) < 1e-13


def test_localbend_concentric_arcs():
    o = O.MDOracle(bent_cilium())
    # 9 bonds x 2 sin(0.05) x (2 * 11 - 10.5 - 10.5)
    assert o.localbend(0) == pytest.approx(0.89962504687221, rel=1e-12)
    o = O.MDOracle(bent_cilium(r=(10.0, 10.5, 10.5)))
    assert o.localbend(0) == pytest.approx(-0.89962504687221, rel=1e-12)


def test_switch_logic_localbend_resets_stroke():
    tstep = 0.05
    # bent beyond max_bend (0.7) in recovery (midpnt > start = 2): back to start
    sw = bent_cilium()
    sw["cilia"][0]["midpnt"] = 5.5
    o = O.MDOracle(sw, globalt=7.0, tstep=tstep)
    o.switch_logic()
    assert o.cilia_state()[0] == (2.0, 7.0)
    # straight: recovery goes on, midpnt += tstep / m_speed
    sw = one_cilium(midpnt=5.5)
    o = O.MDOracle(sw, globalt=7.0, tstep=tstep)
    o.switch_logic()
    mid, ts = o.cilia_state()[0]
    assert ts == 0.0 and mid == pytest.approx(5.5 + tstep / 5.3, abs=1e-15)


def test_switch_logic_recovery_after_100_time_units():
    tstep = 0.05
    sw = one_cilium(midpnt=2.0)
    o = O.MDOracle(sw, globalt=100.0, tstep=tstep)  # globalt - Tstroke = 100: not yet
    o.switch_logic()
    assert o.cilia_state()[0] == (2.0, 0.0)
    o = O.MDOracle(sw, globalt=100.5, tstep=tstep)
    o.switch_logic()
    mid, ts = o.cilia_state()[0]
    assert ts == 0.0 and mid == pytest.approx(3.0 + tstep / 5.3, abs=1e-15)
    # bent filament: the start switch fires, then localbend resets it at once
    sw = bent_cilium()
    sw["cilia"][0]["midpnt"] = 2.0
    o = O.MDOracle(sw, globalt=100.5, tstep=tstep)
    o.switch_logic()
    assert o.cilia_state()[0] == (2.0, 100.5)
