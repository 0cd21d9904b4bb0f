"""ctypes view of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module, and only as the checker (or the timed CPU baseline).
The C restatement lives in ``mpcd_oracle.c``; every function there cites the
reference file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpcd_oracle.so")

RNG_MT, RNG_INJECT, RNG_PHILOX = 0, 1, 2
CALPHA, SALPHA = -0.6427876, 0.76604444  # src/libcilia/mpcd.h:74-75


def build() -> None:
    subprocess.check_call(["sh", os.path.join(HERE, "build_oracle.sh")])


class MT(C.Structure):
    _fields_ = [("s", C.c_uint32 * 624), ("left", C.c_int), ("next", C.c_int)]


class Sphere(C.Structure):
    """Rigid sphere of the oracle (orc_sphere)."""
    _fields_ = [("R", C.c_double), ("Ms", C.c_double), ("Is", C.c_double), ("X", C.c_double * 3),
                ("V", C.c_double * 3), ("W", C.c_double * 3), ("slip", C.c_int), ("dP", C.c_double * 3),
                ("dL", C.c_double * 3), ("inj_sph", C.POINTER(C.c_double))]


def make_sphere(R, Ms, X, V=(0.0, 0.0, 0.0), W=(0.0, 0.0, 0.0), slip=False, Is=None):
    sp = Sphere()
    sp.R, sp.Ms = float(R), float(Ms)
    sp.Is = float(Is) if Is is not None else 0.4 * float(Ms) * float(R) ** 2
    sp.X[:], sp.V[:], sp.W[:] = [float(x) for x in X], [float(x) for x in V], [float(x) for x in W]
    sp.slip = int(bool(slip))
    return sp


class CollideCfg(C.Structure):
    _fields_ = [
        ("L", C.c_int * 3),
        ("calpha", C.c_double),
        ("salpha", C.c_double),
        ("kT", C.c_double),
        ("thermostat", C.c_int),
        ("rng", C.c_int),
        ("mt", C.POINTER(MT)),
        ("inj_gamma", C.POINTER(C.c_double)),
        ("inj_axis", C.POINTER(C.c_double)),
        ("seed", C.c_uint64),
        ("step", C.c_uint64),
        ("rec_gamma", C.POINTER(C.c_double)),
        ("rec_axis", C.POINTER(C.c_double)),
        ("rel_violations", C.c_long),
        ("walls", C.c_int),
        ("rho", C.c_int),
        ("inj_virt", C.POINTER(C.c_double)),
        ("rec_virt", C.POINTER(C.c_double)),
        ("rule", C.c_int),
        ("inj_ran", C.POINTER(C.c_double)),
        ("rec_ran", C.POINTER(C.c_double)),
        ("sphere", C.POINTER(Sphere)),
    ]


SRD, AT, AT_A = 0, 1, 2  # collision rules


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, u32, u64, p = C.c_double, C.c_int, C.c_uint32, C.c_uint64, C.c_void_p
        L.orc_mt_seed.argtypes = [C.POINTER(MT), u32]
        L.orc_mt_u32.argtypes = [C.POINTER(MT)]
        L.orc_mt_u32.restype = u32
        for f in ("orc_mt_rand", "orc_mt_dblexc", "orc_normalrand"):
            getattr(L, f).argtypes = [C.POINTER(MT)]
            getattr(L, f).restype = d
        L.orc_rgamma_mt.argtypes = [C.POINTER(MT), i]
        L.orc_rgamma_mt.restype = d
        L.orc_axis_mt.argtypes = [C.POINTER(MT), p]
        L.orc_rotmat.argtypes = [p, d, d, p]
        L.orc_philox4x64.argtypes = [p, p, p]
        L.orc_philox4x32.argtypes = [p, p, p]
        L.orc_u64_to_open01.argtypes = [u64]
        L.orc_u64_to_open01.restype = d
        L.orc_philox_gamma.argtypes = [u64, u64, u64, i]
        L.orc_philox_gamma.restype = d
        L.orc_philox_axis.argtypes = [u64, u64, u64, p]
        L.orc_philox_shift.argtypes = [u64, u64, p]
        L.orc_box.argtypes = [p, p, p]
        L.orc_box.restype = i
        L.orc_box_many.argtypes = [p, p, p, C.c_long, p]
        L.orc_stream.argtypes = [p, p, C.c_long, d]
        L.orc_stream_walls.argtypes = [p, p, C.c_long, d, i, i, d, d]
        L.orc_philox_virt.argtypes = [u64, u64, u64, i]
        L.orc_philox_virt.restype = d
        L.orc_philox_ran.argtypes = [u64, u64, u64, u64, i]
        L.orc_philox_ran.restype = d
        L.orc_at_omega.argtypes = [p, p, p]
        L.orc_stream_sphere.argtypes = [p, p, C.c_long, d, p, C.POINTER(Sphere)]
        L.orc_sphere_finish.argtypes = [C.POINTER(Sphere), d, p]
        L.orc_philox_sphv.argtypes = [u64, u64, u64, u64, p, p]
        L.orc_fluid_create.argtypes = [C.POINTER(MT), p, C.c_long, p, p]
        L.orc_fluid_create_philox.argtypes = [u64, p, d, C.c_long, p, p]
        L.orc_bin.argtypes = [p, C.c_long, i, p, p, p]
        L.orc_collide.argtypes = [C.POINTER(CollideCfg), p, p, p, C.c_long, p, p, C.c_long]
        L.orc_collide.restype = C.c_longdouble
        L.orc_f32_encode.argtypes = [p, p, p, p, C.c_long, p, p, p]
        L.orc_f32_decode.argtypes = [p, p, p, p, p, C.c_long, p, p]
        L.orc_f32_step.argtypes = [C.POINTER(CollideCfg), p, p, d, p, p, p, C.c_long, p, p, C.c_long, p]
        L.orc_f32_step.restype = C.c_longdouble
        L.orc_md_step.argtypes = [p]
        L.orc_md_recalc.argtypes = [p]
        L.orc_md_switch_logic.argtypes = [p]
        L.orc_md_cilium_length.argtypes = [p, i, i]
        L.orc_md_cilium_length.restype = d
        L.orc_md_refresh_due.argtypes = [d, d]
        L.orc_md_refresh_due.restype = i
        L.orc_md_localbend.argtypes = [p, i]
        L.orc_md_localbend.restype = d
        L.orc_omp_create.argtypes = [p, C.c_long, p, p, i, u32]
        L.orc_omp_create.restype = p
        L.orc_omp_steps.argtypes = [p, i, d]
        L.orc_omp_destroy.argtypes = [p]
        L.orc_omp_threads.argtypes = [p]
        L.orc_omp_threads.restype = i
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i3(L):
    return np.ascontiguousarray(L, dtype=np.int32)


class MTRand:
    """MTRand(uint32 seed) of src/libcilia/MersenneTwister.h."""

    def __init__(self, seed: int):
        self.st = MT()
        lib().orc_mt_seed(C.byref(self.st), seed)

    def u32(self, n=None):
        if n is None:
            return lib().orc_mt_u32(C.byref(self.st))
        return np.array([lib().orc_mt_u32(C.byref(self.st)) for _ in range(n)], dtype=np.uint64)

    def rand(self):
        return lib().orc_mt_rand(C.byref(self.st))

    def dblexc(self):
        return lib().orc_mt_dblexc(C.byref(self.st))

    def normal(self):
        return lib().orc_normalrand(C.byref(self.st))

    def gamma(self, nc):
        return lib().orc_rgamma_mt(C.byref(self.st), nc)

    def axis(self):
        a = np.zeros(3)
        lib().orc_axis_mt(C.byref(self.st), _p(a))
        return a


def rotmat(axis, c=CALPHA, s=SALPHA):
    ax = np.ascontiguousarray(axis, dtype=np.float64)
    R = np.zeros(9)
    lib().orc_rotmat(_p(ax), c, s, _p(R))
    return R.reshape(3, 3)


def philox4x64(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    o = np.zeros(4, dtype=np.uint64)
    lib().orc_philox4x64(_p(c), _p(k), _p(o))
    return o


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32(_p(c), _p(k), _p(o))
    return o


def philox_gamma(seed, step, cell, nc):
    return lib().orc_philox_gamma(seed, step, cell, nc)


def philox_axis(seed, step, cell):
    a = np.zeros(3)
    lib().orc_philox_axis(seed, step, cell, _p(a))
    return a


def philox_shift(seed, step):
    a = np.zeros(3)
    lib().orc_philox_shift(seed, step, _p(a))
    return a


def box(L, a, pos3):
    pos3 = np.ascontiguousarray(pos3, dtype=np.float64).reshape(-1, 3)
    cell = np.zeros(len(pos3), dtype=np.int32)
    lib().orc_box_many(_p(_i3(L)), _p(np.ascontiguousarray(a, dtype=np.float64)), _p(pos3), len(pos3), _p(cell))
    return cell


def stream(pos3, vel4, ts):
    lib().orc_stream(_p(pos3), _p(vel4), len(pos3), ts)


def stream_walls(pos3, vel4, ts, Lz, walls, vw=(0.0, 0.0)):
    """Streaming with the planar-wall bounce-back (walls 1 NOSLIP, 2 SLIP)."""
    lib().orc_stream_walls(_p(pos3), _p(vel4), len(pos3), ts, int(Lz), int(walls), float(vw[0]), float(vw[1]))


def philox_virt(seed, step, cell, k):
    return lib().orc_philox_virt(seed, step, cell, k)


def stream_sphere(pos3, vel4, ts, L, sphere):
    lib().orc_stream_sphere(_p(pos3), _p(vel4), len(pos3), ts, _p(_i3(L)), C.byref(sphere))


def sphere_finish(sphere, ts, L):
    lib().orc_sphere_finish(C.byref(sphere), ts, _p(_i3(L)))


def philox_sphv(seed, step, cell, k):
    u, nr = np.zeros(3), np.zeros(3)
    lib().orc_philox_sphv(seed, step, cell, k, _p(u), _p(nr))
    return u, nr


def philox_ran(seed, step, cell, rank, k):
    return lib().orc_philox_ran(seed, step, cell, rank, k)


def at_omega(I6, L3):
    w = np.zeros(3)
    lib().orc_at_omega(_p(np.ascontiguousarray(I6, dtype=np.float64)), _p(np.ascontiguousarray(L3, dtype=np.float64)),
                       _p(w))
    return w


def fluid_create(mt: MTRand, L, n):
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 4))
    lib().orc_fluid_create(C.byref(mt.st), _p(_i3(L)), n, _p(pos), _p(vel))
    return pos, vel


def fluid_create_philox(seed, ext, n, kT=1.0):
    """Fluid::create (src/libcilia/fluid.cpp:55-98) on the production's
    Philox 'FLUD' stream: (pos3 N x 3, vel4 N x 4)."""
    pos = np.empty((n, 3))
    vel = np.empty((n, 4))
    e = np.asarray(ext, dtype=np.float64)
    lib().orc_fluid_create_philox(int(seed), _p(e), float(kT), int(n), _p(pos), _p(vel))
    return pos, vel


def bin_cells(cell, ncells):
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    start = np.zeros(ncells + 1, dtype=np.int64)
    perm = np.zeros(len(cell), dtype=np.int64)
    rank = np.zeros(len(cell), dtype=np.int32)
    lib().orc_bin(_p(cell), len(cell), ncells, _p(start), _p(perm), _p(rank))
    return start, perm, rank


def make_cfg(L, rng=RNG_INJECT, mt=None, gamma=None, axis=None, seed=0, step=0,
             thermostat=1, kT=1.0, calpha=CALPHA, salpha=SALPHA, record=False, walls=0, rho=10, virt=None,
             rule=SRD, ran=None, n_record_ran=0, sphere=None):
    """Collision configuration.  record=True returns (gamma, axis, virt)
    arrays per cell filled by the collision (NaN where not drawn), plus the
    MPC-AT normals (3 per particle) when n_record_ran > 0."""
    cfg = CollideCfg()
    cfg.L[:] = list(L)
    cfg.calpha, cfg.salpha, cfg.kT = calpha, salpha, kT
    cfg.thermostat, cfg.rng = thermostat, rng
    keep = []
    if mt is not None:
        cfg.mt = C.pointer(mt.st)
    if gamma is not None:
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        keep.append(g)
        cfg.inj_gamma = g.ctypes.data_as(C.POINTER(C.c_double))
    if axis is not None:
        a = np.ascontiguousarray(axis, dtype=np.float64)
        keep.append(a)
        cfg.inj_axis = a.ctypes.data_as(C.POINTER(C.c_double))
    if virt is not None:
        vv = np.ascontiguousarray(virt, dtype=np.float64)
        keep.append(vv)
        cfg.inj_virt = vv.ctypes.data_as(C.POINTER(C.c_double))
    cfg.seed, cfg.step = seed, step
    cfg.walls, cfg.rho = walls, rho
    cfg.rule = rule
    if sphere is not None:
        cfg.sphere = C.pointer(sphere)
        keep.append(sphere)
    if ran is not None:
        rr = np.ascontiguousarray(ran, dtype=np.float64)
        keep.append(rr)
        cfg.inj_ran = rr.ctypes.data_as(C.POINTER(C.c_double))
    ncells = int(np.prod(L))
    rec = None
    if record:
        rec = (np.zeros(ncells), np.zeros(3 * ncells), np.zeros(3 * ncells), np.zeros(3 * max(n_record_ran, 1)))
        keep.extend(rec)
        cfg.rec_gamma = rec[0].ctypes.data_as(C.POINTER(C.c_double))
        cfg.rec_axis = rec[1].ctypes.data_as(C.POINTER(C.c_double))
        cfg.rec_virt = rec[2].ctypes.data_as(C.POINTER(C.c_double))
        if n_record_ran:
            cfg.rec_ran = rec[3].ctypes.data_as(C.POINTER(C.c_double))
    return cfg, keep, rec


def collide(cfg, a, pos3, vel4, bpos3=None, bvel4=None):
    """In-place collision of fluid (pos3, vel4) and beads; returns diss."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    nb = 0 if bpos3 is None else len(bpos3)
    return float(lib().orc_collide(C.byref(cfg), _p(a), _p(pos3), _p(vel4), len(pos3),
                                   _p(bpos3), _p(bvel4), nb))


def f32_encode(L, a, pos3, vel4):
    n = len(pos3)
    cell = np.zeros(n, dtype=np.int32)
    off = np.zeros((n, 3), dtype=np.float32)
    vel = np.zeros((n, 3), dtype=np.float32)
    lib().orc_f32_encode(_p(_i3(L)), _p(np.ascontiguousarray(a, dtype=np.float64)), _p(pos3), _p(vel4),
                         n, _p(cell), _p(off), _p(vel))
    return cell, off, vel


def f32_decode(L, a, cell, off, vel):
    n = len(cell)
    pos = np.zeros((n, 3))
    v4 = np.zeros((n, 4))
    lib().orc_f32_decode(_p(_i3(L)), _p(np.ascontiguousarray(a, dtype=np.float64)), _p(cell), _p(off), _p(vel),
                         n, _p(pos), _p(v4))
    return pos, v4


def f32_step(cfg, a, a_next, ts, cell, off, vel, bpos3=None, bvel4=None, ids=None):
    """One FP32-bulk step in place; ids (int32 per particle) join the
    residual owner's tie-break as the GPU does when ids are tracked."""
    nb = 0 if bpos3 is None else len(bpos3)
    if ids is not None:
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        assert len(ids) == len(cell)
    return float(lib().orc_f32_step(C.byref(cfg), _p(np.ascontiguousarray(a, dtype=np.float64)),
                                     _p(np.ascontiguousarray(a_next, dtype=np.float64)), ts,
                                     _p(cell), _p(off), _p(vel), len(cell), _p(bpos3), _p(bvel4), nb, _p(ids)))


BASELINE_PATH = os.path.join(HERE, "lib", "libmpcd_baseline.so")
_blib = None


def baseline_lib():
    """The same sources built with the reference's flags (-O3 -ffast-math
    -fopenmp, src/Makefile:2), for the timed CPU baseline."""
    global _blib
    if _blib is None:
        if not os.path.exists(BASELINE_PATH):
            build()
        L = C.CDLL(BASELINE_PATH)
        p = C.c_void_p
        L.orc_omp_create.argtypes = [p, C.c_long, p, p, C.c_int, C.c_uint32]
        L.orc_omp_create.restype = p
        L.orc_omp_steps.argtypes = [p, C.c_int, C.c_double]
        L.orc_omp_destroy.argtypes = [p]
        L.orc_omp_threads.argtypes = [p]
        L.orc_omp_threads.restype = C.c_int
        _blib = L
    return _blib


class OmpSim:
    """Threaded bucket algorithm of the reference (src/libcilia/mpcd.cpp:
    140-267, fluid.cpp:152-167): per-thread per-cell pointer buckets,
    static schedules, one MT19937 per thread seeded seed + thread.  baseline
    = True runs the -O3 -ffast-math build (timing); False the oracle's own
    build (checked against orc_collide)."""

    def __init__(self, L, pos3, vel4, threads=0, seed=1, baseline=False):
        self.pos, self.vel = pos3, vel4
        self._L = baseline_lib() if baseline else lib()
        self.h = self._L.orc_omp_create(_p(_i3(L)), len(pos3), _p(pos3), _p(vel4), threads, seed)

    @property
    def threads(self):
        return self._L.orc_omp_threads(self.h)

    def steps(self, k, tstep):
        self._L.orc_omp_steps(self.h, k, tstep)

    def close(self):
        if self.h:
            self._L.orc_omp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# --------------------------------------------------------------- bead MD
class MDSphere(C.Structure):
    _fields_ = [("first", C.c_int), ("npart", C.c_int), ("radius", C.c_double), ("springconstant", C.c_double),
                ("headbond", C.c_double), ("nei_start", C.POINTER(C.c_int)), ("nei", C.POINTER(C.c_int))]


class MDCilium(C.Structure):
    _fields_ = [("first", C.c_int), ("rodlength", C.c_int), ("num_anker", C.c_int), ("constant_beat", C.c_int),
                ("bondlength", C.c_double), ("beatstrength", C.c_double), ("bending", C.c_double),
                ("min_bend", C.c_double), ("max_bend", C.c_double), ("m_speed", C.c_double),
                ("springconstant", C.c_double), ("power_stroke_duration", C.c_double),
                ("recovery_stroke_duration", C.c_double), ("midpnt", C.c_double), ("Tstroke", C.c_double),
                ("newlength", C.POINTER(C.c_double)), ("oldlength", C.POINTER(C.c_double))]


class MDAttach(C.Structure):
    _fields_ = [("bead0", C.c_int), ("bead1", C.c_int), ("l0", C.c_double)]


class MDSystemC(C.Structure):
    _fields_ = [("nb", C.c_int), ("pos3", C.POINTER(C.c_double)), ("vel4", C.POINTER(C.c_double)),
                ("force3", C.POINTER(C.c_double)), ("mobile", C.POINTER(C.c_ubyte)), ("nsph", C.c_int),
                ("sph", C.POINTER(MDSphere)), ("ncil", C.c_int), ("cil", C.POINTER(MDCilium)), ("natt", C.c_int),
                ("att", C.POINTER(MDAttach)), ("att_k", C.c_double), ("globalt", C.c_double),
                ("ststep", C.c_double), ("tstep", C.c_double), ("walls", C.c_int), ("Lz", C.c_int),
                ("vwx", C.c_double), ("vwy", C.c_double)]


def refresh_due(globalt, ststep):
    """Cilia::recalcforce's rest-length refresh test (src/cilia.cpp:206-207)."""
    return bool(lib().orc_md_refresh_due(float(globalt), float(ststep)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class MDOracle:
    """The reference's MD half of a step on a swimmer description (a dict
    as built by tests/md_cases.py).  State arrays are numpy copies."""

    def __init__(self, sw, globalt=0.0, ststep=0.001, tstep=0.05, walls=0, Lz=0, vw=(0.0, 0.0)):
        self.pos = np.ascontiguousarray(sw["pos3"], dtype=np.float64).copy()
        self.vel = np.ascontiguousarray(sw["vel4"], dtype=np.float64).copy()
        self.force = np.zeros_like(self.pos)
        self._keep = []
        sph = (MDSphere * max(1, len(sw["spheres"])))()
        for i, s in enumerate(sw["spheres"]):
            ns = np.ascontiguousarray(s["nei_start"], dtype=np.int32)
            nn = np.ascontiguousarray(s["nei"], dtype=np.int32)
            self._keep += [ns, nn]
            sph[i] = MDSphere(s["first"], s["npart"], s["radius"], s["springconstant"], s["headbond"],
                              ns.ctypes.data_as(C.POINTER(C.c_int)), nn.ctypes.data_as(C.POINTER(C.c_int)))
        self.newlength = [np.array(c["newlength"], dtype=np.float64) for c in sw["cilia"]]
        self.oldlength = [np.array(c["oldlength"], dtype=np.float64) for c in sw["cilia"]]
        cil = (MDCilium * max(1, len(sw["cilia"])))()
        for i, c in enumerate(sw["cilia"]):
            cil[i] = MDCilium(c["first"], c["rodlength"], c["num_anker"], int(c["constant_beat"]), c["bondlength"],
                              c["beatstrength"], c["bending"], c["min_bend"], c["max_bend"], c["m_speed"],
                              c["springconstant"], c["power_stroke_duration"], c["recovery_stroke_duration"],
                              c["midpnt"], c["Tstroke"], _dp(self.newlength[i]), _dp(self.oldlength[i]))
        att = (MDAttach * max(1, len(sw["attach"])))()
        for i, (b0, b1, l0) in enumerate(sw["attach"]):
            att[i] = MDAttach(int(b0), int(b1), float(l0))
        self._sph, self._cil, self._att = sph, cil, att
        self.mobile = None if sw.get("mobile") is None else np.ascontiguousarray(sw["mobile"], dtype=np.uint8)
        self.sys = MDSystemC(len(self.pos), _dp(self.pos), _dp(self.vel), _dp(self.force),
                             None if self.mobile is None else self.mobile.ctypes.data_as(C.POINTER(C.c_ubyte)),
                             len(sw["spheres"]), sph, len(sw["cilia"]), cil, len(sw["attach"]), att,
                             float(sw["att_k"]), float(globalt), float(ststep), float(tstep), int(walls), int(Lz),
                             float(vw[0]), float(vw[1]))

    def step(self):
        lib().orc_md_step(C.byref(self.sys))

    def recalc(self):
        lib().orc_md_recalc(C.byref(self.sys))

    @property
    def globalt(self):
        return self.sys.globalt

    def switch_logic(self):
        lib().orc_md_switch_logic(C.byref(self.sys))

    def cilium_length(self, cil, i):
        return lib().orc_md_cilium_length(C.byref(self.sys), int(cil), int(i))

    def localbend(self, cil):
        return lib().orc_md_localbend(C.byref(self.sys), int(cil))

    def set_cilium(self, cil, **kw):
        """Overwrite scalar fields of cilium cil (midpnt, Tstroke, bending, ...)."""
        for k, v in kw.items():
            setattr(self._cil[cil], k, v)

    def cilia_state(self):
        return [(self._cil[i].midpnt, self._cil[i].Tstroke) for i in range(self.sys.ncil)]
