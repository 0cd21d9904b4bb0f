"""A small swimmer for the MD tests: a bead shell with a centre bead and
cilia anchored on it, in the reference's object layout.

Cilia follow the reference's construction (mcm_amd.swimmer): three rods of
rodlength beads, rotated by random (thetaX, thetaY, thetaZ) and placed
radius - 3 bondlength from the centre (src/cilia.cpp:271-325,
src/clami.cpp:109-134); short stroke durations so that the beat switches
within a few steps.  The attachment
springs follow Clami::add_springs_for_rods_idx (src/clami.cpp:152-178):
for rod beads 0 and 3 of every rod, the closest shell bead and that bead's
first neilist[idx][0] - 1 neighbours.  The shell is the refined icosahedron
of mcm_amd.swimmer (the MD engine takes any neighbour table; the
reference's own refinement and neighbour search, src/sphere.cpp:136-285,
are not on the step's path).
"""
from __future__ import annotations

import numpy as np

from mcm_amd.swimmer import BONDLENGTH, PARTMASS, cilium_positions, icosphere


def shell(levels, radius, center):
    return icosphere(levels, radius, center)


def swimmer(levels=2, radius=4.0, center=(8.0, 8.0, 8.0), n_cilia=6, rodlength=10, seed=0,
            constant_beat=True, beatstrength=0.17, springconstant=20000.0, mobile_anchor=False):
    rng = np.random.default_rng(seed)
    sph_pts, nei_start, nei, headbond = shell(levels, radius, center)
    pos = [sph_pts]
    cilia, attach = [], []
    first = len(sph_pts)
    for c in range(n_cilia):
        tx = rng.uniform(-1.0, 1.0)
        ty = rng.uniform(-1.0, 1.0)
        tz = rng.uniform(0.0, 2.0 * np.pi)
        wo = 3
        r = radius - wo * BONDLENGTH
        base = np.array([r * np.sin(ty), -r * np.cos(ty) * np.sin(tx), r * np.cos(tx) * np.cos(ty)]) + center
        cp = cilium_positions(rodlength, *base, tx, ty, tz)
        pos.append(cp)
        start_delay = rng.uniform(0.0, 0.3)
        cilia.append(dict(first=first, rodlength=rodlength, num_anker=2, constant_beat=constant_beat,
                          bondlength=BONDLENGTH, beatstrength=beatstrength,
                          bending=float(np.sqrt(2 * 5.0 / springconstant)), min_bend=-0.7, max_bend=0.7,
                          m_speed=5.3, springconstant=springconstant, power_stroke_duration=0.15,
                          recovery_stroke_duration=0.3, midpnt=2.0, Tstroke=start_delay if constant_beat else 0.0,
                          newlength=np.full(rodlength, BONDLENGTH), oldlength=np.full(rodlength, BONDLENGTH)))
        for w in (0, wo):
            for j in range(3):
                q = cp[j * rodlength + w]
                dist = ((sph_pts - q) ** 2).sum(axis=1)
                idx = int(np.argmin(dist))
                attach.append((idx, first + j * rodlength + w, float(np.sqrt(dist[idx]))))
                for n in range(nei_start[idx], nei_start[idx + 1] - 1):
                    s = int(nei[n])
                    attach.append((s, first + j * rodlength + w, float(np.sqrt(((sph_pts[s] - q) ** 2).sum()))))
        first += 3 * rodlength
    pos3 = np.vstack(pos)
    vel4 = np.zeros((len(pos3), 4))
    vel4[:, :3] = rng.normal(0.0, 0.05, (len(pos3), 3))
    vel4[:, 3] = PARTMASS
    mobile = None
    if mobile_anchor:  # beads 0 of every rod held fixed (startidx = 1)
        mobile = np.ones(len(pos3), dtype=np.uint8)
        for c in cilia:
            for j in range(3):
                mobile[c["first"] + j * rodlength] = 0
    return dict(pos3=pos3, vel4=vel4, mobile=mobile,
                spheres=[dict(first=0, npart=len(sph_pts), radius=radius, springconstant=springconstant,
                              headbond=headbond, nei_start=nei_start, nei=nei)],
                cilia=cilia, attach=attach, att_k=springconstant)
