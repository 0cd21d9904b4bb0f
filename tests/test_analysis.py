"""Swimmer statistics (mcm_amd/analysis.py, the reference's sr/volvox
analysis) against closed forms: a rigid frame turning at constant angular
velocity, a constant swimming velocity, linear ramps.  No GPU."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from mcm_amd import analysis as A


def _frame(seed=0):
    q = Rotation.random(random_state=seed).as_matrix()
    return q[:, 0], q[:, 1], q[:, 2]


def _series(t, n, b, p, v, power):
    d = {"time": t, "x": 0.01 * t, "y": 0 * t, "z": 0 * t}
    for name, a in (("n", n), ("b", b), ("p", p)):
        for i, c in enumerate("xyz"):
            d[name + c] = a[:, i]
    for i, c in enumerate("xyz"):
        d["v" + c] = v[:, i]
    d["power"] = power
    return d


def test_rotation_rates_of_a_turning_frame():
    """n, b, p turning rigidly at omega: a centred difference at spacing h
    gives sinc(|omega| h) x (omega . axis) exactly, and each axis' own
    projection of omega is constant in time."""
    dt, k, T = 2.5, 12, 12 * 30
    t = np.arange(T) * dt
    omega = np.array([0.004, -0.0025, 0.0015])
    n0, b0, p0 = _frame(3)
    R = Rotation.from_rotvec(np.outer(t, omega))
    n, b, p = R.apply(n0), R.apply(b0), R.apply(p0)
    wn, wb, wp = A.rotation_rates(n, b, p, dt, k)
    h = dt * k
    w = np.linalg.norm(omega)
    sinc = math.sin(w * h) / (w * h)
    inner = slice(k, T - k)
    assert np.abs(wn[inner] - sinc * omega @ n0).max() < 1e-13
    assert np.abs(wb[inner] - sinc * omega @ b0).max() < 1e-13
    assert np.abs(wp[inner] - sinc * omega @ p0).max() < 1e-13
    # one-sided at the two ends of each interleaved series: within (w h)^2 of the rate
    assert np.abs(wn - omega @ n0).max() < (w * h) ** 2


def test_rotation_rates_drop_the_remainder():
    T, k = 103, 10
    n0, b0, p0 = _frame(1)
    n, b, p = (np.tile(x, (T, 1)) for x in (n0, b0, p0))
    wn, wb, wp = A.rotation_rates(n, b, p, 1.0, k)
    assert len(wn) == len(wb) == len(wp) == 100 and np.all(wn == 0.0)


def test_analyze_closed_forms():
    dt, k = 2.5, 6
    t = np.arange(400) * dt
    n0, b0, p0 = _frame(5)
    w = 0.001  # n turns about b: phi(t) = w (t - t_first)
    R = Rotation.from_rotvec(np.outer(t, w * b0))
    n, b, p = R.apply(n0), R.apply(b0), R.apply(p0)
    V = np.tile([-1.0e-3, 2.0e-3, 3.0e-3], (len(t), 1))
    power = np.arange(len(t), dtype=float)
    data = _series(t, 2.0 * n, 3.0 * b, 0.5 * p, V, power)  # unnormalised axes, as the .san holds
    eq = 500.0
    r, nn, vv = A.analyze(data, eq_time=eq, k=k)
    sel = t >= eq
    ts = t[sel] - t[sel][0]
    assert sel.sum() == len(nn) == 200
    assert abs(r["phi_mean"] - w * ts.mean()) < 1e-12 and abs(r["phi_max"] - w * ts[-1]) < 1e-12
    assert abs(r["phi_std"] - w * np.std(ts, ddof=1)) < 1e-12
    assert (r["vx_m"], r["vy_m"], r["vz_m"]) == pytest.approx((-1e-3, 2e-3, 3e-3), abs=1e-18)
    # the reference's quirk: the y and z "abs" entries hold the x component's
    assert r["v_abs_x_m"] == r["v_abs_y_m"] == r["v_abs_z_m"] == pytest.approx(1e-3, abs=1e-18)
    vn = nn @ np.array([-1.0e-3, 2.0e-3, 3.0e-3])
    assert abs(r["v_normal_m"] - vn.mean()) < 1e-15
    assert r["e_kin_m"] == pytest.approx(math.sqrt(14.0) * 1e-3, rel=1e-14) and r["e_kin_std"] < 1e-15
    m = np.arange(200, 400, dtype=float)
    assert r["power_total"] == m.sum() and r["power_mean"] == m.mean()
    assert r["power_std"] == pytest.approx(math.sqrt(200 * 201 / 12.0), rel=1e-12)  # ddof = 1 of a ramp
    h = dt * k
    assert r["Omega_b_mean"] == pytest.approx(w * math.sin(w * h) / (w * h), rel=0.01)
    assert abs(r["Omega_n_mean"]) < 1e-6 and abs(r["Omega_p_mean"]) < 1e-6


def test_load_san_layouts(tmp_path):
    row17 = " ".join(str(x) for x in range(17))
    (tmp_path / "a.san").write_text(row17 + "\n" + row17 + "\n")
    d = A.load_san(tmp_path / "a.san")
    assert list(d["time"]) == [0.0, 0.0] and d["power"][0] == 16.0 and np.isnan(d["ekin"]).all()
    row11 = " ".join(str(x) for x in range(11))
    (tmp_path / "b.san").write_text(row11 + "\n")
    d = A.load_san(tmp_path / "b.san")
    assert d["nz"][0] == 9.0 and d["power"][0] == 10.0 and np.isnan(d["bx"]).all()
    row18 = " ".join(str(x) for x in range(18))
    (tmp_path / "c.san").write_text(row18 + "\n")
    d = A.load_san(tmp_path / "c.san")
    assert d["ekin"][0] == 7.0 and d["nx"][0] == 8.0 and d["power"][0] == 17.0
    (tmp_path / "d.san").write_text("1 2 3 4 5\n")
    with pytest.raises(ValueError):
        A.load_san(tmp_path / "d.san")
