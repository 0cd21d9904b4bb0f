"""Writes volvox_sim.cfg: the configuration file the reference's generator
(create.py:55-111, run as its __main__ at create.py:121-122: equator_num 10,
k_theta 1, energy 1.0, thetaZ 0, phase_lack = -pi + pi/7) writes for the
paper's swimmer, in libconfig syntax as pylibconfig2 emits it.  The values
follow create.py's formulas: rings at -45, 0, 45 degrees holding
round(10 cos 45) = 7, 10 and 7 cilia, theta offsets pi/10, 0, pi/10,
start delays (k_phi phi + k_theta theta mod 2 pi) / 2 pi x 225, rounded to
2 decimals (create.py:43-52, 94-101), bending sqrt(2 E / k) = 0.01.

Run: python tests/golden/make_volvox_cfg.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def delays(phi_rings, per_ring, k_theta, offsets, phase_lag, taub):
    k_phi = phase_lag * 4 / np.pi
    out = []
    for idx, phi in enumerate(phi_rings):
        for theta in np.arange(per_ring[idx]) / per_ring[idx] * 2 * np.pi + offsets[idx]:
            out.append(k_phi * phi + k_theta * theta)
    return np.round(np.array(out) % (2 * np.pi) / 2 / np.pi * taub, 2)


def main():
    eq = 10
    ring2 = int(np.round(eq * np.cos(45.0 / 180 * np.pi)))
    per_ring = [ring2, eq, ring2]
    y_thetas = [-45.0, 0.0, 45.0]
    delta = 2 * np.pi / eq / 2
    offsets = [delta, 0.0, delta]
    phase_lag = (np.linspace(-np.pi, np.pi, 7, endpoint=False) + 2 * np.pi / 14)[0]
    taub = 75.0 + 150.0
    d = delays(np.array(y_thetas) / 180 * np.pi, per_ring, 1, offsets, phase_lag, taub)
    f = lambda xs: "[ " + ", ".join(repr(float(x)) for x in xs) + " ]"
    text = f"""name = "volvox";
globalt = 0.0;
tstep = 0.05;
runtime = 85680.0;
sizex = 100;
sizey = 100;
sizez = 100;
omega = 0.0;
desiredt = {float(np.ceil(1000 * taub))!r};
md_objects = ( {{
  name = "volvox";
  cilia = {{
    bending = {float(np.sqrt(2 * 1.0 / 20000.0))!r};
    springconstant = 20000.0;
    beatstrength = 0.17;
    angle = 0.0;
    name = "ConstantBeatCilia";
    rodlength = 26;
    thetaZ = 0.0;
    thetaX = 0.0;
    thetaY = 0.0;
    m_speed = 5.3;
    max_bend = 0.7;
    position = [ 0.0, 0.0, 0.0 ];
    min_bend = -0.7;
    write_data = False;
    power_stroke_duration = 75.0;
    recovery_stroke_duration = 150.0;
    start_delay = 0.0;
    type = "ConstantBeatCilia";
  }};
  position = [ 0.0, 0.0, 0.0 ];
  type = "Clami";
  springconstant = 20000.0;
  radius = 8.0;
  y_thetas = {f(y_thetas)};
  num_equators = [ {", ".join(str(n) for n in per_ring)} ];
  theta_offset = {f(offsets)};
  start_delays = {f(d)};
}} );
"""
    with open(os.path.join(HERE, "volvox_sim.cfg"), "w") as fp:
        fp.write(text)


if __name__ == "__main__":
    main()
