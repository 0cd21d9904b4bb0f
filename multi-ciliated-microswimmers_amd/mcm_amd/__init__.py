"""Python view of the gfx950 MPCD library (C ABI in include/mpcd_abi.h).

The library must be built (``build_lib.py``); importing this module on a
machine without it raises instead of falling back to anything else.  The
class names follow the reference's solvent interface: ``MPCD`` holds the
cell grid, the random shift and the solvent (src/libcilia/mpcd.h:19-40 and
src/libcilia/fluid.h:16-40 fused, since the device owns the particle arrays).
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libmcm_mpcd.so")
HEADER = os.path.join(os.path.dirname(PKG_DIR), "include", "mpcd_abi.h")

FP64, FP32BULK = 0, 1
WALLS_NONE, NOSLIP, SLIP = 0, 1, 2  # planar walls at z = 0 and z = size[2] - 1
SRD, AT, AT_A = 0, 1, 2  # collision rules: the reference's SRD; MPC-AT without / with angular momentum
CALPHA, SALPHA = -0.6427876, 0.76604444  # src/libcilia/mpcd.h:74-75


class MPCDError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("size", C.c_int * 3),
        ("rho", C.c_int),
        ("kT", C.c_double),
        ("calpha", C.c_double),
        ("salpha", C.c_double),
        ("precision", C.c_int),
        ("thermostat", C.c_int),
        ("track_ids", C.c_int),
        ("seed", C.c_ulonglong),
        ("tile_variant", C.c_int),
        ("walls", C.c_int),
        ("wall_v", C.c_double * 2),
        ("collision", C.c_int),
    ]


class Draws(C.Structure):
    _fields_ = [("gamma", C.c_void_p), ("axis", C.c_void_p), ("virt", C.c_void_p), ("ran", C.c_void_p),
                ("sph", C.c_void_p)]


class SphereParams(C.Structure):
    _fields_ = [("radius", C.c_double), ("mass", C.c_double), ("inertia", C.c_double), ("x", C.c_double * 3),
                ("v", C.c_double * 3), ("w", C.c_double * 3), ("slip", C.c_int)]


class Diag(C.Structure):
    _fields_ = [
        ("momentum", C.c_double * 3),
        ("twice_kinetic", C.c_double),
        ("diss_last", C.c_double),
        ("diss_total", C.c_double),
        ("momentum_q", C.c_longlong * 3),
        ("n", C.c_longlong),
        ("rel_violations", C.c_longlong),
        ("flags", C.c_uint),
        ("deferred_tiles", C.c_longlong),
        ("max_tile_load", C.c_longlong),
        ("tile_capacity", C.c_longlong),
    ]


class MDSphere(C.Structure):
    _fields_ = [("first", C.c_int), ("npart", C.c_int), ("nei_off", C.c_int), ("radius", C.c_double),
                ("springconstant", C.c_double), ("headbond", C.c_double)]


class MDCilium(C.Structure):
    _fields_ = [("first", C.c_int), ("rodlength", C.c_int), ("num_anker", C.c_int), ("constant_beat", C.c_int),
                ("bondlength", C.c_double), ("beatstrength", C.c_double), ("bending", C.c_double),
                ("min_bend", C.c_double), ("max_bend", C.c_double), ("m_speed", C.c_double),
                ("springconstant", C.c_double), ("power_stroke_duration", C.c_double),
                ("recovery_stroke_duration", C.c_double), ("midpnt", C.c_double), ("Tstroke", C.c_double)]


class MDAttachment(C.Structure):
    _fields_ = [("bead0", C.c_int), ("bead1", C.c_int), ("l0", C.c_double)]


class MDDesc(C.Structure):
    _fields_ = [("nb", C.c_int), ("pos3", C.c_void_p), ("vel4", C.c_void_p), ("mobile", C.c_void_p),
                ("nsph", C.c_int), ("sph", C.POINTER(MDSphere)), ("nei_start", C.c_void_p), ("nei", C.c_void_p),
                ("ncil", C.c_int), ("cil", C.POINTER(MDCilium)), ("newlength", C.c_void_p),
                ("oldlength", C.c_void_p), ("natt", C.c_int), ("att", C.POINTER(MDAttachment)),
                ("att_springconstant", C.c_double), ("globalt", C.c_double), ("ststep", C.c_double)]


_lib = None


def declared_symbols():
    """Entry points declared in include/mpcd_abi.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mpcd_[a-z_0-9]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MPCDError(f"HIP library not built: {LIB_PATH} (run build_lib.py)")
        L = C.CDLL(LIB_PATH)
        p, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
        L.mpcd_version.restype = C.c_char_p
        L.mpcd_last_error.restype = C.c_char_p
        L.mpcd_create.argtypes = [C.POINTER(Params), C.POINTER(p)]
        for f in ("mpcd_destroy", "mpcd_sync", "mpcd_reset_diss"):
            getattr(L, f).argtypes = [p]
        L.mpcd_set_shift.argtypes = [p, p]
        L.mpcd_get_shift.argtypes = [p, p]
        L.mpcd_draw_shift.argtypes = [p, p]
        L.mpcd_upload_fluid.argtypes = [p, p, p, ll]
        L.mpcd_download_fluid.argtypes = [p, p, p, ll]
        L.mpcd_create_fluid.argtypes = [p, ll]
        L.mpcd_count.argtypes = [p]
        L.mpcd_count.restype = ll
        L.mpcd_stage_beads.argtypes = [p, p, p, i]
        L.mpcd_fetch_beads.argtypes = [p, p, i]
        L.mpcd_step.argtypes = [p, d, p, p, p]
        L.mpcd_step_injected.argtypes = [p, d, p, C.POINTER(Draws)]
        L.mpcd_sphere_attach.argtypes = [p, C.POINTER(SphereParams)]
        L.mpcd_sphere_detach.argtypes = [p]
        L.mpcd_sphere_state.argtypes = [p, C.POINTER(SphereParams), p, p]
        L.mpcd_diagnostics.argtypes = [p, C.POINTER(Diag)]
        L.mpcd_arm_debug.argtypes = [p, i]
        L.mpcd_debug_binning.argtypes = [p, p, p, ll]
        L.mpcd_set_timing.argtypes = [p, i]
        L.mpcd_timing.argtypes = [p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.mpcd_tile_variants.argtypes = [i, p, i]
        L.mpcd_autotune.argtypes = [C.POINTER(Params), ll, i, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.mpcd_md_attach.argtypes = [p, C.POINTER(MDDesc)]
        L.mpcd_md_detach.argtypes = [p]
        L.mpcd_md_run.argtypes = [p, d]
        L.mpcd_md_fetch.argtypes = [p, p, p, p, p, p, p, C.POINTER(C.c_double)]
        L.mpcd_checkpoint_save.argtypes = [p, C.c_char_p]
        L.mpcd_checkpoint_load.argtypes = [p, C.c_char_p]
        L.mpcd_ensemble_create.argtypes = [p, i, p]
        L.mpcd_ensemble_step.argtypes = [p, d, i]
        L.mpcd_ensemble_set_graph.argtypes = [p, i]
        L.mpcd_ensemble_set_md_wide.argtypes = [p, i]
        L.mpcd_ensemble_set_md_block.argtypes = [p, i]
        L.mpcd_ensemble_autotune.argtypes = [p, d, i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
        for f in ("mpcd_ensemble_sync", "mpcd_ensemble_destroy"):
            getattr(L, f).argtypes = [p]
        _lib = L
    return _lib


def make_params(size, rho=10, kT=1.0, calpha=CALPHA, salpha=SALPHA, precision=FP64,
                thermostat=True, track_ids=False, seed=1, tile_variant=-1, walls=WALLS_NONE, wall_v=(0.0, 0.0),
                collision=SRD):
    prm = Params()
    prm.size[:] = [int(s) for s in size]
    prm.rho, prm.kT, prm.calpha, prm.salpha = int(rho), float(kT), float(calpha), float(salpha)
    prm.precision, prm.thermostat, prm.track_ids = int(precision), int(bool(thermostat)), int(bool(track_ids))
    prm.seed, prm.tile_variant = int(seed), int(tile_variant)
    prm.walls = int(walls)
    prm.wall_v[:] = [float(wall_v[0]), float(wall_v[1])]
    prm.collision = int(collision)
    return prm


def tile_variants(precision):
    """(tx, ty, tz, particles per thread) of each step-kernel variant."""
    buf = np.zeros(64, dtype=np.int32)
    n = lib().mpcd_tile_variants(int(precision), buf.ctypes.data_as(C.c_void_p), 16)
    _chk(0 if n >= 0 else n)
    return [tuple(int(x) for x in buf[4 * i:4 * i + 4]) for i in range(n)]


def autotune(size, n, steps=5, **kw):
    """Times every tile variant on a scratch context; returns (variant, ms)."""
    prm = make_params(size, **kw)
    best, ms = C.c_int(), C.c_double()
    _chk(lib().mpcd_autotune(C.byref(prm), int(n), int(steps), C.byref(best), C.byref(ms)))
    return best.value, ms.value


def _chk(rc):
    if rc != 0:
        raise MPCDError(f"status {rc}: {lib().mpcd_last_error().decode()}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, cols=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if cols is not None:
        a = a.reshape(-1, cols)
    return a


class MPCD:
    """One MPCD solvent on the GPU: MPCD(int sizex, int sizey, int sizez)."""

    def __init__(self, size, rho=10, kT=1.0, calpha=CALPHA, salpha=SALPHA, precision=FP64,
                 thermostat=True, track_ids=False, seed=1, tile_variant=-1, walls=WALLS_NONE, wall_v=(0.0, 0.0),
                 collision=SRD):
        prm = make_params(size, rho, kT, calpha, salpha, precision, thermostat, track_ids, seed, tile_variant,
                          walls, wall_v, collision)
        self.walls = int(walls)
        self.collision = int(collision)
        self.size = tuple(prm.size)
        self.ncells = int(np.prod(self.size))
        self.precision = precision
        h = C.c_void_p()
        _chk(lib().mpcd_create(C.byref(prm), C.byref(h)))
        self.h = h

    # -- lifetime
    def close(self):
        ens = getattr(self, "_ens", None)
        if ens is not None:
            ens.close()
        if getattr(self, "h", None):
            lib().mpcd_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- grid shift (mpcd.h:56, mpcd.cpp:249-253)
    def set_shift(self, a):
        _chk(lib().mpcd_set_shift(self.h, _p(_f64(a))))

    def get_shift(self):
        a = np.zeros(3)
        _chk(lib().mpcd_get_shift(self.h, _p(a)))
        return a

    def draw_shift(self):
        a = np.zeros(3)
        _chk(lib().mpcd_draw_shift(self.h, _p(a)))
        return a

    # -- solvent (fluid.h)
    def upload_fluid(self, pos3, vel4):
        pos3, vel4 = _f64(pos3, 3), _f64(vel4, 4)
        assert len(pos3) == len(vel4)
        _chk(lib().mpcd_upload_fluid(self.h, _p(pos3), _p(vel4), len(pos3)))

    def create_fluid(self, n):
        _chk(lib().mpcd_create_fluid(self.h, int(n)))

    @property
    def n(self):
        return lib().mpcd_count(self.h)

    def download_fluid(self):
        n = self.n
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 4))
        _chk(lib().mpcd_download_fluid(self.h, _p(pos), _p(vel), n))
        return pos, vel

    # -- beads (mdobject.cpp:62-72, 153-160)
    def stage_beads(self, pos3, vel4):
        pos3, vel4 = _f64(pos3, 3), _f64(vel4, 4)
        self._nb = len(pos3)
        _chk(lib().mpcd_stage_beads(self.h, _p(pos3), _p(vel4), len(pos3)))

    def fetch_beads(self, nb):
        v = np.zeros((nb, 4))
        _chk(lib().mpcd_fetch_beads(self.h, _p(v), nb))
        return v

    # -- one step: collide + stream + rebin (mpcd.cpp:140-247, fluid.cpp:152-167)
    def step(self, tstep, next_shift=None, inj_gamma=None, inj_axis=None, inj_virt=None, inj_ran=None,
             inj_sph=None):
        """Injected draws (mpcd_draws): inj_virt the virtual wall particles'
        normals (3 per cell); inj_ran the MPC-AT normals (3 per particle id,
        beads after the solvent); inj_sph the sphere's virtual-particle
        normals (3 per candidate, rho per cell)."""
        ns = None if next_shift is None else _f64(next_shift)
        arrs = {k: (None if x is None else _f64(x)) for k, x in
                dict(gamma=inj_gamma, axis=inj_axis, virt=inj_virt, ran=inj_ran, sph=inj_sph).items()}
        if arrs["gamma"] is not None:
            assert arrs["gamma"].size == self.ncells and arrs["axis"].size == 3 * self.ncells
        if arrs["virt"] is not None:
            assert arrs["virt"].size == 3 * self.ncells
        if any(x is not None for k, x in arrs.items() if k not in ("gamma", "axis")) or self.walls or \
                self.collision != SRD or getattr(self, "_sphere", False):
            d = Draws(*[None if arrs[k] is None else arrs[k].ctypes.data for k in ("gamma", "axis", "virt", "ran", "sph")])
            _chk(lib().mpcd_step_injected(self.h, float(tstep), _p(ns), C.byref(d)))
        else:
            _chk(lib().mpcd_step(self.h, float(tstep), _p(ns), _p(arrs["gamma"]), _p(arrs["axis"])))

    # -- rigid sphere (not in the reference)
    def sphere_attach(self, radius, mass, x, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0), slip=False, inertia=0.0):
        sp = SphereParams(float(radius), float(mass), float(inertia), (C.c_double * 3)(*x), (C.c_double * 3)(*v),
                          (C.c_double * 3)(*w), int(bool(slip)))
        _chk(lib().mpcd_sphere_attach(self.h, C.byref(sp)))
        self._sphere = True

    def sphere_detach(self):
        _chk(lib().mpcd_sphere_detach(self.h))
        self._sphere = False

    def sphere_state(self):
        sp = SphereParams()
        f, t = np.zeros(3), np.zeros(3)
        _chk(lib().mpcd_sphere_state(self.h, C.byref(sp), _p(f), _p(t)))
        return dict(x=np.array(sp.x[:]), v=np.array(sp.v[:]), w=np.array(sp.w[:]), radius=sp.radius, mass=sp.mass,
                    inertia=sp.inertia, force=f, torque=t)

    def diagnostics(self):
        d = Diag()
        _chk(lib().mpcd_diagnostics(self.h, C.byref(d)))
        return {
            "momentum": np.array(d.momentum[:]),
            "twice_kinetic": d.twice_kinetic,
            "diss_last": d.diss_last,
            "diss_total": d.diss_total,
            "momentum_q": np.array(d.momentum_q[:], dtype=np.int64),
            "n": d.n,
            "rel_violations": d.rel_violations,
            "flags": d.flags,
            "deferred_tiles": d.deferred_tiles,
            "max_tile_load": d.max_tile_load,
            "tile_capacity": d.tile_capacity,
        }

    def reset_diss(self):
        _chk(lib().mpcd_reset_diss(self.h))

    def arm_debug(self, on=True):
        _chk(lib().mpcd_arm_debug(self.h, int(on)))

    def debug_binning(self):
        n = self.n
        cell = np.zeros(n, dtype=np.int32)
        rank = np.zeros(n, dtype=np.int32)
        _chk(lib().mpcd_debug_binning(self.h, _p(cell), _p(rank), n))
        return cell, rank

    def set_timing(self, on=True):
        _chk(lib().mpcd_set_timing(self.h, int(on)))

    def timing(self):
        ms, n = C.c_double(), C.c_longlong()
        _chk(lib().mpcd_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def sync(self):
        _chk(lib().mpcd_sync(self.h))

    # -- attached swimmer (bead MD, src/simulation.cpp:217-265)
    def md_attach(self, sw, globalt=0.0, ststep=0.001):
        """Hands a swimmer (a dict: pos3, vel4, mobile, spheres, cilia,
        attach, att_k) to the device; mpcd_step then collides its beads and
        runs its MD sub-steps."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        sph = (MDSphere * max(1, len(sw["spheres"])))()
        nei_start, nei, off = [], [], 0
        for i, sp in enumerate(sw["spheres"]):
            sph[i] = MDSphere(sp["first"], sp["npart"], off, sp["radius"], sp["springconstant"], sp["headbond"])
            nei_start.extend(int(x) + len(nei) for x in sp["nei_start"])
            nei.extend(int(x) for x in sp["nei"])
            off += sp["npart"] + 1
        cil = (MDCilium * max(1, len(sw["cilia"])))()
        nl, ol = [], []
        for i, c in enumerate(sw["cilia"]):
            cil[i] = MDCilium(c["first"], c["rodlength"], c["num_anker"], int(c["constant_beat"]), c["bondlength"],
                              c["beatstrength"], c["bending"], c["min_bend"], c["max_bend"], c["m_speed"],
                              c["springconstant"], c["power_stroke_duration"], c["recovery_stroke_duration"],
                              c["midpnt"], c["Tstroke"])
            nl.extend(c["newlength"])
            ol.extend(c["oldlength"])
        att = (MDAttachment * max(1, len(sw["attach"])))()
        for i, (b0, b1, l0) in enumerate(sw["attach"]):
            att[i] = MDAttachment(int(b0), int(b1), float(l0))
        nb = len(sw["pos3"])
        desc = MDDesc(nb, arr(sw["pos3"], np.float64), arr(sw["vel4"], np.float64),
                      None if sw.get("mobile") is None else arr(sw["mobile"], np.uint8),
                      len(sw["spheres"]), sph, arr(nei_start or [0], np.int32), arr(nei or [0], np.int32),
                      len(sw["cilia"]), cil, arr(nl or [0.0], np.float64), arr(ol or [0.0], np.float64),
                      len(sw["attach"]), att, float(sw["att_k"]), float(globalt), float(ststep))
        _chk(lib().mpcd_md_attach(self.h, C.byref(desc)))
        self._md = (nb, len(sw["cilia"]), len(nl))

    # -- checkpoints (fluid_n.dat, <name>.dat, mpcd.dat restarts)
    def checkpoint_save(self, path):
        _chk(lib().mpcd_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        _chk(lib().mpcd_checkpoint_load(self.h, os.fsencode(path)))

    def md_detach(self):
        _chk(lib().mpcd_md_detach(self.h))
        self._md = None

    def md_run(self, tstep):
        _chk(lib().mpcd_md_run(self.h, float(tstep)))

    def md_fetch(self):
        nb, ncil, nlen = self._md
        pos, vel, frc = np.zeros((nb, 3)), np.zeros((nb, 4)), np.zeros((nb, 3))
        beat, nl, ol = np.zeros((max(ncil, 1), 2)), np.zeros(max(nlen, 1)), np.zeros(max(nlen, 1))
        g = C.c_double()
        _chk(lib().mpcd_md_fetch(self.h, _p(pos), _p(vel), _p(frc), _p(beat), _p(nl), _p(ol), C.byref(g)))
        return dict(pos=pos, vel=vel, force=frc, beat=beat[:ncil], newlength=nl[:nlen], oldlength=ol[:nlen],
                    globalt=g.value)


class Ensemble:
    """Independent boxes (MPCD contexts, each with its own seed and optional
    swimmer) stepped together: per step one launch per phase for all boxes,
    replayed from a captured HIP graph.  Each box's step is bit-identical to
    stepping it alone.  While in the ensemble the boxes refuse state changes;
    downloads and diagnostics work after sync()."""

    def __init__(self, boxes, graph=True, md_wide=True):
        self.boxes = list(boxes)
        arr = (C.c_void_p * len(self.boxes))(*[b.h.value for b in self.boxes])
        h = C.c_void_p()
        _chk(lib().mpcd_ensemble_create(arr, len(self.boxes), C.byref(h)))
        self.h = h
        for b in self.boxes:
            b._ens = self
        if not graph:
            self.set_graph(False)
        if not md_wide:
            self.set_md_wide(False)

    def set_graph(self, on=True):
        _chk(lib().mpcd_ensemble_set_graph(self.h, int(bool(on))))

    def set_md_wide(self, on=True):
        """MD half: one launch per phase over all swimmers (True) or a
        workgroup per swimmer (False); the same results."""
        _chk(lib().mpcd_ensemble_set_md_wide(self.h, int(bool(on))))

    def set_md_block(self, threads):
        """Threads per workgroup of the wide MD launches (256, 512, 1024)."""
        _chk(lib().mpcd_ensemble_set_md_block(self.h, int(threads)))

    def autotune(self, tstep, steps=2):
        """Times the MD half's launch shapes on this ensemble's own steps
        (4 (steps + 1) steps of the run) and keeps the fastest.  Returns
        {"md_wide", "md_block", "ms_per_step", "candidates_ms"}; the
        candidates are the wide path at 256, 512, 1024 threads per workgroup
        and a workgroup per swimmer."""
        w, b, ms = C.c_int(), C.c_int(), C.c_double()
        each = (C.c_double * 4)()
        _chk(lib().mpcd_ensemble_autotune(self.h, float(tstep), int(steps), C.byref(w), C.byref(b), C.byref(ms),
                                          each))
        return {"md_wide": bool(w.value), "md_block": b.value, "ms_per_step": ms.value,
                "candidates_ms": {"wide_256": each[0], "wide_512": each[1], "wide_1024": each[2],
                                  "workgroup_per_swimmer": each[3]}}

    def step(self, tstep, nsteps=1):
        _chk(lib().mpcd_ensemble_step(self.h, float(tstep), int(nsteps)))

    def sync(self):
        _chk(lib().mpcd_ensemble_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().mpcd_ensemble_destroy(self.h)
            self.h = None
            for b in self.boxes:
                b._ens = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
