"""Python view of the configuration-file driver (C ABI in include/mcm_sim.h).

``Simulation(path)`` reads a configuration file of the reference's `cilia`
executable (src/main.cpp:40-100) and builds its objects on the host
(Simulation::load_config / add_md_object, src/simulation.cpp:266-351);
``start`` creates the solvent and attaches the objects on the GPU, ``run``
is Simulation::run (src/simulation.cpp:59-158).  The bead model the
factory built is readable before the start (``model()``), with no GPU.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import FP64, MDDesc, MPCDError, lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "include",
                      "mcm_sim.h")
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "cilia")


class SimInfo(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("size", C.c_int * 3), ("globalt", C.c_double), ("desiredt", C.c_double),
                ("tstep", C.c_double), ("runtime", C.c_double), ("omega", C.c_double), ("nobjects", C.c_int),
                ("nbeads", C.c_int), ("nspheres", C.c_int), ("ncilia", C.c_int), ("nattach", C.c_int),
                ("steps", C.c_longlong)]


class SimObject(C.Structure):
    _fields_ = [("type", C.c_char * 32), ("name", C.c_char * 128), ("write_data", C.c_int), ("first", C.c_int),
                ("nbeads", C.c_int), ("sphere", C.c_int), ("ncilia", C.c_int), ("first_cilium", C.c_int),
                ("first_attach", C.c_int), ("nattach", C.c_int)]


class SimOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("seed", C.c_ulonglong), ("measureinterval", C.c_int),
                ("max_steps", C.c_longlong), ("restart", C.c_int), ("quiet", C.c_int), ("ref_files", C.c_int)]


_bound = False


def declared_symbols():
    """Entry points declared in include/mcm_sim.h."""
    return sorted(set(re.findall(r"\b(mcm_sim_[a-z_0-9]+)\s*\(", open(HEADER).read())))


def _lib():
    global _bound
    L = lib()
    if not _bound:
        p, i = C.c_void_p, C.c_int
        L.mcm_sim_last_error.restype = C.c_char_p
        L.mcm_sim_load.argtypes = [C.c_char_p, C.POINTER(p)]
        L.mcm_sim_free.argtypes = [p]
        L.mcm_sim_get_info.argtypes = [p, C.POINTER(SimInfo)]
        L.mcm_sim_get_object.argtypes = [p, i, C.POINTER(SimObject)]
        L.mcm_sim_md_desc.argtypes = [p, C.POINTER(MDDesc)]
        L.mcm_sim_start.argtypes = [p, C.POINTER(SimOpts)]
        L.mcm_sim_next_tstep.argtypes = [p, i]
        for f in ("mcm_sim_writedata", "mcm_sim_run", "mcm_sim_save"):
            getattr(L, f).argtypes = [p]
        L.mcm_sim_run_ensemble.argtypes = [C.POINTER(p), i]
        L.mcm_sim_context.argtypes = [p]
        L.mcm_sim_context.restype = p
        _bound = True
    return L


def _chk(rc):
    if rc != 0:
        raise MPCDError(_lib().mcm_sim_last_error().decode())


def _arr(ptr, n, dtype, shape=None):
    if n == 0:
        return np.zeros(shape if shape else 0, dtype=dtype)
    ct = {np.float64: C.c_double, np.int32: C.c_int}[dtype]
    a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()
    return a.reshape(shape) if shape else a


class Simulation:
    """A configuration file's simulation (see the module docstring)."""

    def __init__(self, path):
        self.h = C.c_void_p()
        _chk(_lib().mcm_sim_load(os.fsencode(path), C.byref(self.h)))

    def close(self):
        if self.h:
            _lib().mcm_sim_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = SimInfo()
        _chk(_lib().mcm_sim_get_info(self.h, C.byref(i)))
        return dict(name=i.name.decode(), size=tuple(i.size), globalt=i.globalt, desiredt=i.desiredt, tstep=i.tstep,
                    runtime=i.runtime, omega=i.omega, nobjects=i.nobjects, nbeads=i.nbeads, nspheres=i.nspheres,
                    ncilia=i.ncilia, nattach=i.nattach, steps=i.steps)

    def objects(self):
        out = []
        for k in range(self.info()["nobjects"]):
            o = SimObject()
            _chk(_lib().mcm_sim_get_object(self.h, k, C.byref(o)))
            out.append(dict(type=o.type.decode(), name=o.name.decode(), write_data=bool(o.write_data),
                            first=o.first, nbeads=o.nbeads, sphere=o.sphere, ncilia=o.ncilia,
                            first_cilium=o.first_cilium, first_attach=o.first_attach, nattach=o.nattach))
        return out

    def model(self):
        """The bead model handed to mpcd_md_attach, as the swimmer dict of
        ``MPCD.md_attach`` (pos3, vel4, spheres with local CSR neighbours,
        cilia, attach, att_k) plus globalt and ststep."""
        d = MDDesc()
        _chk(_lib().mcm_sim_md_desc(self.h, C.byref(d)))
        nb = d.nb
        pos3 = _arr(d.pos3, 3 * nb, np.float64, (nb, 3))
        vel4 = _arr(d.vel4, 4 * nb, np.float64, (nb, 4))
        spheres = []
        nst = sum(d.sph[i].npart + 1 for i in range(d.nsph))
        nei_start_all = _arr(d.nei_start, nst, np.int32)
        nnei = int(nei_start_all[-1]) if nst else 0
        nei_all = _arr(d.nei, nnei, np.int32)
        for i in range(d.nsph):
            s = d.sph[i]
            ns = nei_start_all[s.nei_off:s.nei_off + s.npart + 1]
            spheres.append(dict(first=s.first, npart=s.npart, radius=s.radius, springconstant=s.springconstant,
                                headbond=s.headbond, nei_start=(ns - ns[0]).astype(np.int32),
                                nei=nei_all[ns[0]:ns[-1]].copy()))
        cilia, off = [], 0
        nl_total = sum(d.cil[i].rodlength for i in range(d.ncil))
        nl = _arr(d.newlength, nl_total, np.float64)
        ol = _arr(d.oldlength, nl_total, np.float64)
        for i in range(d.ncil):
            c = d.cil[i]
            cilia.append(dict(first=c.first, rodlength=c.rodlength, num_anker=c.num_anker,
                              constant_beat=bool(c.constant_beat), bondlength=c.bondlength,
                              beatstrength=c.beatstrength, bending=c.bending, min_bend=c.min_bend,
                              max_bend=c.max_bend, m_speed=c.m_speed, springconstant=c.springconstant,
                              power_stroke_duration=c.power_stroke_duration,
                              recovery_stroke_duration=c.recovery_stroke_duration, midpnt=c.midpnt, Tstroke=c.Tstroke,
                              newlength=nl[off:off + c.rodlength].copy(), oldlength=ol[off:off + c.rodlength].copy()))
            off += c.rodlength
        attach = [(d.att[i].bead0, d.att[i].bead1, d.att[i].l0) for i in range(d.natt)]
        return dict(pos3=pos3, vel4=vel4, mobile=None, spheres=spheres, cilia=cilia, attach=attach,
                    att_k=d.att_springconstant, globalt=d.globalt, ststep=d.ststep)

    def start(self, precision=FP64, seed=0, measureinterval=50, max_steps=0, restart=True, quiet=True,
              ref_files=False):
        o = SimOpts(precision, seed, measureinterval, max_steps, int(restart), int(quiet), int(ref_files))
        _chk(_lib().mcm_sim_start(self.h, C.byref(o)))

    def next_tstep(self, n=1):
        _chk(_lib().mcm_sim_next_tstep(self.h, int(n)))

    def writedata(self):
        _chk(_lib().mcm_sim_writedata(self.h))

    def run(self):
        _chk(_lib().mcm_sim_run(self.h))

    def save(self):
        _chk(_lib().mcm_sim_save(self.h))

    def context(self):
        """The MPCD context handle (an ``mpcd_ctx *``)."""
        return _lib().mcm_sim_context(self.h)


def run_ensemble(sims):
    arr = (C.c_void_p * len(sims))(*[s.h.value for s in sims])
    _chk(_lib().mcm_sim_run_ensemble(arr, len(sims)))
