"""Solvent observables: the quantities the long-run statistics are judged
by, measured on an ``MPCD`` context.

* ``debug``: the totals of MPCD::debug (src/libcilia/mpcd.cpp:404-467):
  particle count, mass, momentum per mass, sum m v^2 / 3 / count (kT) and
  the dissipation total, over the solvent and optionally the beads;
* ``velocity_moments``: kT per axis and the kurtosis of each velocity
  component (3 for a Maxwell-Boltzmann distribution);
* ``shear_wave`` / ``shear_wave_viscosity``: the kinematic viscosity from
  the decay of a transverse wave v_x = A sin(k z), exp(-nu k^2 t);
* ``srd_viscosity``: the kinetic plus collisional viscosity of SRD with
  random shifts (Kikuchi, Pooley, Ryder, Yeomans, J. Chem. Phys. 119 (2003)
  6388, eqs. (35)-(36); a = m = kT = 1, M particles per cell, rotation
  angle arccos(calpha)).
"""
from __future__ import annotations

import numpy as np


def debug(m, bead_vel4=None):
    """MPCD::debug's totals over the solvent (downloaded) and, if given,
    the beads (N x 4, mass in column 3)."""
    _, v = m.download_fluid()
    if bead_vel4 is not None:
        v = np.vstack([v, np.asarray(bead_vel4, dtype=np.float64)])
    mass = v[:, 3]
    M = float(mass.sum())
    mom = (v[:, :3] * mass[:, None]).sum(axis=0)
    e2 = float((mass * (v[:, :3] ** 2).sum(axis=1)).sum())
    d = m.diagnostics()
    return dict(num=len(v), mass=M, momentum_per_mass=mom / M, kT=e2 / 3.0 / len(v), diss_total=d["diss_total"])


def velocity_moments(vel):
    """(kT per axis, kurtosis per axis) of unit-mass velocities."""
    w = np.asarray(vel)[:, :3]
    w = w - w.mean(axis=0)
    var = (w ** 2).mean(axis=0)
    return var, (w ** 4).mean(axis=0) / var ** 2


def srd_viscosity(M, h, calpha):
    a = np.arccos(calpha)
    e = np.exp(-M)
    nu_kin = h / 2.0 * (5.0 * M / ((M - 1 + e) * (2.0 - np.cos(a) - np.cos(2 * a))) - 1.0)
    nu_coll = (1.0 - np.cos(a)) * (M - 1 + e) / (18.0 * h * M)
    return float(nu_kin + nu_coll)


def shear_wave(L, rho, A, rng, kT=1.0):
    """Uniform positions and Maxwell-Boltzmann velocities (zero mean) with
    the wave v_x += A sin(2 pi z / L_z) added: (pos N x 3, vel N x 4, k)."""
    n = rho * int(np.prod(L))
    k = 2.0 * np.pi / L[2]
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L, dtype=np.float64)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(scale=np.sqrt(kT), size=(n, 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    vel[:, 0] += A * np.sin(k * pos[:, 2])
    return pos, vel, k


def wave_amplitude(m, k):
    p, v = m.download_fluid()
    return 2.0 * float(np.mean(v[:, 0] * np.sin(k * p[:, 2])))


def shear_wave_viscosity(m, k, tstep, steps, every):
    """Steps the context `steps` times, sampling the wave amplitude every
    `every` steps; returns (nu, times, amplitudes) from a fit of
    log(amplitude) against time."""
    amps, times = [], []
    for s in range(steps + 1):
        if s % every == 0:
            amps.append(wave_amplitude(m, k))
            times.append(s * tstep)
        if s < steps:
            m.step(tstep)
    amps, times = np.array(amps), np.array(times)
    slope = np.polyfit(times, np.log(amps), 1)[0]
    return -slope / (k * k), times, amps
