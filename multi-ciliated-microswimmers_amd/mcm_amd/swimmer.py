"""Host-side construction of a Clami-type swimmer for ``MPCD.md_attach``.

A swimmer is a bead shell with a centre bead (the reference's Sphere,
src/sphere.cpp:12-41) carrying 3-rod cilia (src/cilia.cpp:271-325) held by
attachment springs (src/clami.cpp:152-178).  The dict returned here is the
object description the device engine takes: bead arrays (every object's
beads in one array, each object contiguous and rod-major), the sphere with
its neighbour table, the cilia with their parameters and beat state, and
the attachment springs in Clami's order.

The shell is a refined icosahedron (12, 42, 162 or 642 surface beads, the
reference's shell sizes 163 and 643 with the centre); its neighbours are the
mesh edges and ``headbond`` their mean length.  Cilia are anchored at given
directions (by default quasi-uniform Fibonacci sites); each is placed
radius - 3 bondlength from the centre and rotated so that its rods point
outward, as Clami::create_cilia does (src/clami.cpp:109-134).
"""
from __future__ import annotations

import numpy as np

BONDLENGTH = 0.5  # src/cilia.h:78
PARTMASS = 5.0  # src/libcilia/mdobject.h:76


def icosphere(levels, radius, center):
    """Centre bead plus a refined icosahedron: (positions, nei_start, nei,
    headbond) with the neighbour table in CSR form over local bead index
    (bead 0, the centre, has none)."""
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
    pts = np.vstack([np.zeros(3), np.array(v) * radius]) + np.asarray(center, dtype=np.float64)
    nei_start, nei = [0, 0], []
    for i in range(n_surface):
        nei.extend(sorted(j + 1 for j in adj[i]))
        nei_start.append(len(nei))
    edges = [np.linalg.norm(pts[i + 1] - pts[j]) for i in range(n_surface)
             for j in nei[nei_start[i + 1]:nei_start[i + 2]]]
    return pts, np.array(nei_start, dtype=np.int32), np.array(nei, dtype=np.int32), float(np.mean(edges))


def cilium_positions(R, wx, wy, wz, tx, ty, tz):
    """Cilia::create (src/cilia.cpp:271-325): three rods of R beads along z,
    rotated by (thetaX, thetaY, thetaZ), moved to (wx, wy, wz)."""
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


def fibonacci_directions(n):
    k = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * k
    return np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)


def clami(levels=3, radius=5.0, center=(16.0, 16.0, 16.0), n_cilia=16, rodlength=10, directions=None,
          seed=0, constant_beat=True, beatstrength=0.17, springconstant=20000.0, energy=5.0,
          power_stroke_duration=75.0, recovery_stroke_duration=150.0, m_speed=5.3, max_bend=0.7, min_bend=-0.7,
          start_delays=None, thermal_velocity=0.0):
    """Swimmer description (see the module docstring).  Cilium parameters
    default to the reference's generator (create.py: k = 20000, bending =
    sqrt(2 E / k), beatstrength 0.17, m_speed 5.3, bend limits 0.7, power
    and recovery stroke 75 and 150 time units); num_anker is 2 (the
    reference's Setting constructor, src/cilia.h:11-16)."""
    if rodlength <= 3:
        raise ValueError("rodlength must exceed 3: cilia attach to the shell at beads 0 and 3 (src/clami.cpp:124)")
    rng = np.random.default_rng(seed)
    center = np.asarray(center, dtype=np.float64)
    sph_pts, nei_start, nei, headbond = icosphere(levels, radius, center)
    dirs = fibonacci_directions(n_cilia) if directions is None else np.asarray(directions, dtype=np.float64)
    if start_delays is None:
        start_delays = rng.uniform(0.0, power_stroke_duration + recovery_stroke_duration, len(dirs))
    pos = [sph_pts]
    cilia, attach = [], []
    first = len(sph_pts)
    wo = 3  # attachment depth in bonds (src/clami.cpp:124)
    r = radius - wo * BONDLENGTH
    for c, d in enumerate(dirs):
        d = d / np.linalg.norm(d)
        ty = float(np.arcsin(np.clip(d[0], -1.0, 1.0)))
        tx = float(np.arctan2(-d[1], d[2]))
        tz = float(rng.uniform(0.0, 2.0 * np.pi))
        base = np.array([r * np.sin(ty), -r * np.cos(ty) * np.sin(tx), r * np.cos(tx) * np.cos(ty)]) + center
        cp = cilium_positions(rodlength, *base, tx, ty, tz)
        pos.append(cp)
        cilia.append(dict(first=first, rodlength=rodlength, num_anker=2, constant_beat=constant_beat,
                          bondlength=BONDLENGTH, beatstrength=beatstrength,
                          bending=float(np.sqrt(2 * energy / springconstant)), min_bend=min_bend, max_bend=max_bend,
                          m_speed=m_speed, springconstant=springconstant,
                          power_stroke_duration=power_stroke_duration,
                          recovery_stroke_duration=recovery_stroke_duration, midpnt=2.0,
                          Tstroke=float(start_delays[c]) if constant_beat else 0.0,
                          newlength=np.full(rodlength, BONDLENGTH), oldlength=np.full(rodlength, BONDLENGTH)))
        # Clami::add_springs_for_rods_idx for rod beads 0 and wo: the closest
        # shell bead and its first neilist[idx][0] - 1 neighbours
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
    if thermal_velocity > 0:
        vel4[:, :3] = rng.normal(0.0, thermal_velocity, (len(pos3), 3))
    vel4[:, 3] = PARTMASS
    return dict(pos3=pos3, vel4=vel4, mobile=None,
                spheres=[dict(first=0, npart=len(sph_pts), radius=radius, springconstant=springconstant,
                              headbond=headbond, nei_start=nei_start, nei=nei)],
                cilia=cilia, attach=attach, att_k=springconstant)
