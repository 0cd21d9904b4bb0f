"""A small swimmer for the MD tests: a bead shell with a centre bead and
cilia anchored on it, in the reference's object layout.

Cilia follow the reference's construction: three rods of rodlength beads,
rotated by (thetaX, thetaY, thetaZ) and placed radius - 3 bondlength from
the centre (src/cilia.cpp:271-325, src/clami.cpp:109-134).  The attachment
springs follow Clami::add_springs_for_rods_idx (src/clami.cpp:152-178):
for rod beads 0 and 3 of every rod, the closest shell bead and that bead's
first neilist[idx][0] - 1 neighbours.  The shell is a refined icosahedron
built here (the MD engine takes any neighbour table; the reference's own
refinement and neighbour search, src/sphere.cpp:136-285, are not on the
step's path).
"""
from __future__ import annotations

import numpy as np

BONDLENGTH = 0.5  # src/cilia.h:78
PARTMASS = 5.0  # src/libcilia/mdobject.h:76


def shell(levels, radius, center):
    """Centre bead plus a refined icosahedron on the sphere (12, 42, 162 or
    642 surface beads for levels 0-3, the reference's shell sizes 163 and
    643 with the centre, src/sphere.cpp:18-27); neighbours are the mesh
    edges (5 or 6 per bead), headbond their mean length."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(x, dtype=np.float64) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(levels):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[key] = len(v) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    n_surface = len(v)
    adj = [set() for _ in range(n_surface)]
    for a, b, c in f:
        adj[a] |= {b, c}
        adj[b] |= {a, c}
        adj[c] |= {a, b}
    pts = np.vstack([np.zeros(3), np.array(v) * radius]) + np.asarray(center)
    nei_start, nei = [0, 0], []
    for i in range(n_surface):
        nei.extend(sorted(j + 1 for j in adj[i]))
        nei_start.append(len(nei))
    edges = [np.linalg.norm(pts[i + 1] - pts[j]) for i in range(n_surface) for j in nei[nei_start[i + 1]:nei_start[i + 2]]]
    return pts, np.array(nei_start, dtype=np.int32), np.array(nei, dtype=np.int32), float(np.mean(edges))


def cilium_positions(R, wx, wy, wz, tx, ty, tz):
    """Cilia::create (src/cilia.cpp:271-325)."""
    bl = BONDLENGTH
    l0 = bl * np.sin(np.pi / 3.0) * 2.0 / 3.0
    p = np.zeros((3, R, 3))
    i = np.arange(R)
    p[0, :, 2] = bl * i
    p[0, :, 0] -= l0
    p[1, :, 2] = bl * i
    p[1, :, 1] = bl / 2.0
    p[1, :, 0] = bl * np.sin(np.pi / 3.0) - l0
    p[2, :, 2] = bl * i
    p[2, :, 1] = -bl / 2.0
    p[2, :, 0] = bl * np.sin(np.pi / 3.0) - l0
    x, y, z = p[..., 0].copy(), p[..., 1].copy(), p[..., 2].copy()
    out = np.empty_like(p)
    out[..., 0] = z * np.sin(ty) + np.cos(ty) * (x * np.cos(tz) - y * np.sin(tz))
    out[..., 1] = np.cos(tx) * (y * np.cos(tz) + x * np.sin(tz)) - np.sin(tx) * (
        z * np.cos(ty) + np.sin(ty) * (-(x * np.cos(tz)) + y * np.sin(tz)))
    out[..., 2] = np.sin(tx) * (y * np.cos(tz) + x * np.sin(tz)) + np.cos(tx) * (
        z * np.cos(ty) + np.sin(ty) * (-(x * np.cos(tz)) + y * np.sin(tz)))
    out += np.array([wx, wy, wz])
    return out.reshape(3 * R, 3)


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
