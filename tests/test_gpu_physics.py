"""Physics of long runs on the GPU: FP32-bulk against FP64 at fixed error.

The two forms draw different random numbers, so their trajectories part
after the first step; they must agree in what the physics fixes:

* shear viscosity: a transverse shear wave v_x = A sin(k z) decays as
  exp(-nu k^2 t).  nu from both forms must agree with each other and with
  the kinetic + collisional SRD viscosity of the random-shift method
  (Kikuchi, Pooley, Ryder, Yeomans, J. Chem. Phys. 119 (2003) 6388, eqs.
  (35)-(36) with rho = M particles per cell, a = m = kT = 1):
      nu_kin  = h/2 [5M / ((M - 1 + e^-M)(2 - cos a - cos 2a)) - 1]
      nu_coll = (1 - cos a) (M - 1 + e^-M) / (18 h M)
  (the measurement and the formula: mcm_amd/observables.py)
* the velocity distribution: from a uniform (non-Maxwellian) start both
  forms reach the Maxwell-Boltzmann kurtosis 3 and kT = 1;
* drift: over 10^5 steps, with SRD and with MPC-AT+a, the FP32-bulk
  momentum is exactly conserved (an integer) and the FP64 momentum drifts
  by less than 1e-10 per particle.
"""
import numpy as np
import pytest

from mcm_amd import observables as OBS

pytestmark = pytest.mark.gpu

H = 0.1
CALPHA = -0.6427876  # the reference's truncated constants (mpcd.h:74-75)
SALPHA = 0.76604444


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _shear_wave_nu(mcm, prec, pos, vel, L, k, steps, every, seed):
    with mcm.MPCD(L, precision=prec, seed=seed) as m:
        m.upload_fluid(pos, vel)
        nu, _, amps = OBS.shear_wave_viscosity(m, k, H, steps, every)
        d = m.diagnostics()
    assert d["flags"] == 0
    return nu, amps


def test_shear_viscosity_fp32_fp64_and_theory(mcm):
    L = (32, 32, 32)
    _print_case("shear_viscosity", L=L)
    A = 0.5
    pos, vel, k = OBS.shear_wave(L, 10, A, np.random.default_rng(77))
    nu32, a32 = _shear_wave_nu(mcm, mcm.FP32BULK, pos, vel, L, k, 300, 10, seed=5)
    nu64, a64 = _shear_wave_nu(mcm, mcm.FP64, pos, vel, L, k, 300, 10, seed=6)
    nu_th = OBS.srd_viscosity(10, H, CALPHA)
    print(f"[result] nu fp32 {nu32:.4f} fp64 {nu64:.4f} theory {nu_th:.4f}", flush=True)
    assert abs(a32[0] - A) < 0.02 and abs(a64[0] - A) < 0.02
    assert abs(nu32 / nu64 - 1.0) < 0.06
    assert abs(nu32 / nu_th - 1.0) < 0.1
    assert abs(nu64 / nu_th - 1.0) < 0.1


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_maxwell_boltzmann_from_uniform_start(mcm, prec_name):
    L = (32, 32, 32)
    _print_case("maxwell_boltzmann", prec=prec_name, L=L)
    rng = np.random.default_rng(3)
    n = 10 * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.uniform(-np.sqrt(3.0), np.sqrt(3.0), (n, 3))  # variance 1, kurtosis 1.8
    vel[:, :3] -= vel[:, :3].mean(0)
    with mcm.MPCD(L, precision=getattr(mcm, prec_name), seed=9) as m:
        m.upload_fluid(pos, vel)
        for _ in range(30):
            m.step(H)
        _, v = m.download_fluid()
        d = m.diagnostics()
    assert d["flags"] == 0
    var, kurt = OBS.velocity_moments(v)
    print(f"[result] {prec_name} kT {var} kurtosis {kurt}", flush=True)
    assert np.all(np.abs(var - 1.0) < 0.01)
    assert np.all(np.abs(kurt - 3.0) < 0.04)  # sigma ~ sqrt(96/n) = 0.017


@pytest.mark.parametrize("prec_name,rule_name", [("FP32BULK", "SRD"), ("FP64", "SRD"), ("FP32BULK", "AT_A"),
                                                 ("FP64", "AT_A")])
def test_momentum_drift_1e5_steps(mcm, prec_name, rule_name):
    L = (16, 16, 16)
    steps = 100000
    _print_case("momentum_drift", prec=prec_name, rule=rule_name, L=L, steps=steps)
    n = 10 * int(np.prod(L))
    with mcm.MPCD(L, precision=getattr(mcm, prec_name), seed=21, collision=getattr(mcm, rule_name)) as m:
        m.create_fluid(n)
        m.step(H)
        d0 = m.diagnostics()
        for _ in range(steps):
            m.step(H)
        d1 = m.diagnostics()
    assert d0["flags"] == 0 and d1["flags"] == 0
    assert d1["n"] == n
    if prec_name == "FP32BULK":
        assert np.array_equal(d1["momentum_q"], d0["momentum_q"])  # exact, an integer
    else:
        drift = np.abs(np.asarray(d1["momentum"]) - np.asarray(d0["momentum"])).max() / n
        print(f"[result] FP64 {rule_name} momentum drift per particle over {steps} steps: {drift:.3e}", flush=True)
        assert drift < 1e-10
    assert abs(d1["twice_kinetic"] / (3.0 * n) - 1.0) < 0.02


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_energy_without_thermostat(mcm, prec_name):
    """Without the thermostat an SRD rotation keeps each cell's kinetic
    energy only if cos^2 + sin^2 = 1.  With exact constants the total
    changes by rounding alone (FP64: roundoff; FP32-bulk: its 2^-21 grid and
    FP32 map, a random walk).  The reference's truncated constants
    (mpcd.h:74-75) give cos^2 + sin^2 = 1 - 1.743e-8, which scales the
    velocity components normal to the axis: per step the total changes by
    (cos^2 + sin^2 - 1) x 2/3 x (relative share of the energy, 1 - cells/N
    = 0.9 at 10 per cell).  Both forms must show that change."""
    L = (32, 32, 32)
    steps = 200
    n = 10 * int(np.prod(L))
    a = np.deg2rad(130.0)
    res = {}
    for name, c, s_ in (("exact", np.cos(a), np.sin(a)), ("truncated", CALPHA, SALPHA)):
        _print_case("energy_no_thermostat", prec=prec_name, constants=name, steps=steps)
        with mcm.MPCD(L, precision=getattr(mcm, prec_name), seed=12, thermostat=False, calpha=c, salpha=s_) as m:
            m.create_fluid(n)
            m.step(H)
            e0 = m.diagnostics()["twice_kinetic"]
            for _ in range(steps):
                m.step(H)
            d = m.diagnostics()
        assert d["flags"] == 0
        res[name] = (d["twice_kinetic"] - e0) / e0
    predicted = (1.0 + (CALPHA ** 2 + SALPHA ** 2 - 1.0) * (2.0 / 3.0) * 0.9) ** steps - 1.0
    print(f"[result] {prec_name} relative energy change over {steps} steps without thermostat: exact constants "
          f"{res['exact']:.3e}, truncated {res['truncated']:.4e} (predicted {predicted:.4e})", flush=True)
    assert abs(res["exact"]) < (1e-12 if prec_name == "FP64" else 1e-7)
    assert abs(res["truncated"] / predicted - 1.0) < 0.02
