"""Swimmer statistics from a Clami's ``<name>.san`` time series.

A numpy restatement of the reference's analysis (sr/volvox/io.py:59-84
``load_data``, sr/volvox/process.py:3-76 ``get_rotation_vector_ts_p``,
``normalize``, ``get_orientation_vec``, ``analyze``), which process.py runs
over every simulation directory:

* orientation: the shell axes n, b, p (to beads 1, 32, 137), normalised;
* phi: the angle between n and its value at the first sample after the
  equilibration time;
* the centre-of-mass velocity and its component along n (the swimming
  speed ``v_normal``);
* rotation rates Omega_n = (p x dp/dt).n, Omega_b = (p x dp/dt).b,
  Omega_p = (n x dn/dt).p from numpy gradients of every k-th sample
  (k = 240, k interleaved series);
* the speed |v| (named ``e_kin`` there) and the dissipated power column.

Statistics are pandas' (skip NaN; standard deviations with ddof = 1), and
the keys are the reference's, including its ``v_abs_y_*`` and ``v_abs_z_*``
entries, which hold the x component's values (process.py:47-53).  The
reference asserts a sample spacing of exactly 2.5 time units
(process.py:37-38); here the spacing is taken from the data (mean of the
time differences, as there) unless given.
"""
from __future__ import annotations

import numpy as np

SAN_COLUMNS = ("time", "x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "bx", "by", "bz", "px", "py", "pz",
               "power")
_V1_COLUMNS = ("time", "x", "y", "z", "vx", "vy", "vz", "nx", "ny", "nz", "power")
_EKIN_COLUMNS = ("time", "x", "y", "z", "vx", "vy", "vz", "ekin", "nx", "ny", "nz", "bx", "by", "bz", "px", "py",
                 "pz", "power")


def load_san(path):
    """Columns of a .san file as a dict of arrays (load_data,
    sr/volvox/io.py:59-84): 17 columns as Clami::writedata writes them,
    the 11-column first version, or 18 with an ``ekin`` column; columns a
    version lacks are NaN."""
    rows = [l.split() for l in open(path) if l.strip()]
    if not rows:
        raise ValueError(f"{path}: no samples")
    ncol = len(rows[0])
    names = {17: SAN_COLUMNS, 11: _V1_COLUMNS, 18: _EKIN_COLUMNS}.get(ncol)
    if names is None or any(len(r) != ncol for r in rows):
        raise ValueError(f"{path}: rows of {ncol} columns are not a .san layout")
    a = np.array(rows, dtype=np.float64)
    out = {n: a[:, i] for i, n in enumerate(names)}
    for n in set(_EKIN_COLUMNS) - set(names):
        out[n] = np.full(len(a), np.nan)
    return out


def _unit(v):
    return v / np.sqrt((v ** 2).sum(axis=1))[:, None]


def orientation(data):
    """n, b, p normalised (get_orientation_vec)."""
    return tuple(_unit(np.stack([data[c + "x"], data[c + "y"], data[c + "z"]], axis=1)) for c in "nbp")


def rotation_rates(n, b, p, dt, k=1):
    """Omega_n, Omega_b, Omega_p (get_rotation_vector_ts_p): for each of the k
    interleaved series of every k-th sample (the last len % k samples
    dropped), the cross product of each axis with its numpy gradient at
    spacing dt k, projected as above.  Returned in time order."""
    m = len(n) - len(n) % k
    idx, wn, wb, wp = [], [], [], []
    for l in range(k):
        sel = np.arange(l, m, k)
        e_, b_, p_ = n[sel], b[sel], p[sel]
        ie = np.cross(e_, np.gradient(e_, dt * k, axis=0))
        ip = np.cross(p_, np.gradient(p_, dt * k, axis=0))
        wn.append((ip * e_).sum(axis=1))
        wb.append((ip * b_).sum(axis=1))
        wp.append((ie * p_).sum(axis=1))
        idx.append(sel)
    order = np.argsort(np.concatenate(idx), kind="stable")
    return tuple(np.concatenate(w)[order] for w in (wn, wb, wp))


def _mean(x):
    return float(np.nanmean(x))


def _std(x):
    x = np.asarray(x)[~np.isnan(x)]
    return float(np.std(x, ddof=1)) if len(x) > 1 else float("nan")


def analyze(data, eq_time=500.0, dt=None, k=240):
    """The reference's ``analyze``: (results dict, n, v) over the samples
    at time >= eq_time."""
    t = np.asarray(data["time"])
    if dt is None:
        dt = float(np.mean(np.diff(t)))
    keep = t >= eq_time
    d = {c: np.asarray(v)[keep] for c, v in data.items()}
    n, b, p = orientation(d)
    r = {}
    phi = np.arccos((n * n[0]).sum(axis=1))
    r["phi_mean"], r["phi_std"], r["phi_max"] = _mean(phi), _std(phi), float(np.nanmax(phi))
    v = np.stack([d["vx"], d["vy"], d["vz"]], axis=1)
    r["vx_m"], r["vy_m"], r["vz_m"] = _mean(v[:, 0]), _mean(v[:, 1]), _mean(v[:, 2])
    ax = np.abs(v[:, 0])
    for c in "xyz":  # each from the x component, as the reference's
        r[f"v_abs_{c}_m"] = _mean(ax)
    for c in "xyz":
        r[f"v_abs_{c}_std"] = _std(ax)
    vn = (n * v).sum(axis=1)
    r["v_normal_m"], r["v_normal_std"] = _mean(vn), _std(vn)
    wn, wb, wp = rotation_rates(n, b, p, dt, k)
    r["Omega_n_mean"], r["Omega_n_std"] = _mean(wn), _std(wn)
    r["Omega_b_mean"], r["Omega_b_std"] = _mean(wb), _std(wb)
    r["Omega_p_mean"], r["Omega_p_std"] = _mean(wp), _std(wp)
    speed = np.sqrt((v * v).sum(axis=1))
    r["e_kin_m"], r["e_kin_std"] = _mean(speed), _std(speed)
    pw = d["power"]
    r["power_total"] = float(np.nansum(pw))
    r["power_mean"], r["power_std"] = _mean(pw), _std(pw)
    return r, n, v
