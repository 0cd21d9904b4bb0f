"""The rigid sphere in the FP32-bulk form (not in the reference), on the GPU.

The FP32-bulk form takes the sphere as the FP64 form does (the same virtual
particles from the same draws, the same bounce-back); it stores a bounced
particle's velocity back on the 2^-21 grid and gives the sphere the fluid's
residual sum in the cells it cuts.  The FP64 form is bit-exact against the
oracle (test_gpu_sphere.py), so the FP32-bulk form is held to it:

* one step from the same state with the same injected draws (Gamma
  variates, axes, the sphere's virtual-particle normals): every particle's
  position and velocity within 1e-5, the sphere's velocity, angular
  velocity, force and torque within 1e-5 relative;
* 50 steps on a 32^3 box: fluid + sphere momentum conserved to FP64
  roundoff (the fluid's as exact integers), no particle inside the sphere;
* Stokes' drag with Hasimoto's correction, as for the FP64 form.
"""
import numpy as np
import pytest

from cases import random_draws
from test_gpu_sphere import _fluid_outside, _inside, _print_case

pytestmark = pytest.mark.gpu

H = 0.1
Q = 2.0 ** 21


def _on_grid(vel):
    v = vel.copy()
    v[:, :3] = np.rint(v[:, :3] * Q) / Q
    return v


def _min_image(d, L):
    L = np.asarray(L, dtype=float)
    return d - L * np.rint(d / L)


@pytest.mark.parametrize("slip", [False, True])
def test_sphere_step_fp32_against_fp64(mcm, slip):
    L = (16, 12, 12)
    X, R = (8.3, 5.7, 6.2), 3.1
    V, W = (0.05, -0.03, 0.02), (0.1, -0.2, 0.15)
    Ms = 4.0 / 3.0 * np.pi * R ** 3 * 10
    _print_case("sphere_step_f32", slip=slip, L=L)
    rng = np.random.default_rng(43)
    pos, vel = _fluid_outside(L, X, R, rng)
    vel = _on_grid(vel)
    ncells = int(np.prod(L))
    a = rng.uniform(0, 1, 3)
    a_next = rng.uniform(0, 1, 3)
    g, ax = random_draws(ncells, rng)
    sph = rng.normal(size=3 * 10 * ncells)
    out = {}
    for form, prec in (("fp64", mcm.FP64), ("fp32", mcm.FP32BULK)):
        with mcm.MPCD(L, precision=prec, track_ids=True, seed=12) as m:
            m.set_shift(a)
            m.upload_fluid(pos, vel)
            m.sphere_attach(R, Ms, X, V, W, slip=slip)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax, inj_sph=sph)
            gp, gv = m.download_fluid()
            out[form] = (gp, gv, m.sphere_state(), m.diagnostics()["flags"])
    p64, v64, s64, f64 = out["fp64"]
    p32, v32, s32, f32 = out["fp32"]
    assert f64 == 0 and f32 == 0
    dp = np.abs(_min_image(p32 - p64, L)).max()
    dv = np.abs(v32[:, :3] - v64[:, :3]).max()
    print(f"[result] slip={slip}: max |dx| {dp:.2e}, max |dv| {dv:.2e}", flush=True)
    assert dp < 1e-5 and dv < 1e-5
    for key in ("v", "w", "force", "torque"):
        a64, a32 = np.asarray(s64[key]), np.asarray(s32[key])
        err = np.abs(a32 - a64).max() / max(np.abs(a64).max(), 1e-30)
        print(f"[result]   sphere {key}: {a64} vs {a32}, relative {err:.2e}", flush=True)
        assert err < 1e-5, key
    np.testing.assert_allclose(s32["x"], s64["x"], rtol=0, atol=1e-12)


def test_sphere_fp32_invariants(mcm):
    L = (32, 32, 32)
    X, R = (16.0, 15.5, 16.5), 5.0
    Ms = 4.0 / 3.0 * np.pi * R ** 3 * 10
    _print_case("sphere_f32_invariants", L=L)
    rng = np.random.default_rng(9)
    pos, vel = _fluid_outside(L, X, R, rng)
    vel = _on_grid(vel)
    with mcm.MPCD(L, precision=mcm.FP32BULK, seed=3) as m:
        m.upload_fluid(pos, vel)
        m.sphere_attach(R, Ms, X, v=(0.2, 0.0, 0.0))
        P0 = vel[:, :3].sum(0) + Ms * np.array([0.2, 0.0, 0.0])
        for _ in range(50):
            m.step(H)
        st = m.sphere_state()
        d = m.diagnostics()
        assert d["flags"] == 0
        P1 = d["momentum"] + Ms * st["v"]
        print(f"[result] momentum drift {np.abs(P1 - P0).max():.2e}", flush=True)
        np.testing.assert_allclose(P1, P0, rtol=0, atol=1e-8)
        gp, gv = m.download_fluid()
        np.testing.assert_allclose(gv[:, :3].sum(0), d["momentum"], rtol=0, atol=1e-9)
        assert _inside(gp, st["x"], R, L).sum() == 0
        assert st["v"][0] < 0.2  # the solvent drags the sphere


def test_sphere_fp32_refusals(mcm):
    with mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK, tile_variant=1) as m:
        with pytest.raises(mcm.MPCDError):
            m.sphere_attach(3.0, 100.0, (8.0, 8.0, 8.0))
    with mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK) as m:
        m.sphere_attach(3.0, 100.0, (8.0, 8.0, 8.0))
        with pytest.raises(mcm.MPCDError):
            m.stage_beads(np.full((1, 3), 2.0), np.array([[0.0, 0.0, 0.0, 10.0]]))


def test_stokes_drag_fp32(mcm):
    """The FP64 form's Stokes-drag measurement (test_gpu_sphere.py) in the
    FP32-bulk form: radius 4, V = 0.1, periodic 48^3, 6000 steps after 150,
    the mean force against 6 pi eta R (V - U) with Hasimoto's correction;
    within 0.8 .. 1.1 of it, as the FP64 form."""
    from mcm_amd import observables as OBS
    L, R, V = 48, 4.0, 0.1
    _print_case("stokes_drag_f32", L=L, R=R)
    rng = np.random.default_rng(5)
    n = 10 * L ** 3
    pos = rng.uniform(0, 1, (n, 3)) * L
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    c = np.array([L / 2.0] * 3)
    keep = ((pos - c) ** 2).sum(1) > R * R
    pos, vel = pos[keep], _on_grid(vel[keep])
    F, U = [], []
    with mcm.MPCD((L, L, L), precision=mcm.FP32BULK, seed=3) as m:
        m.upload_fluid(pos, vel)
        m.sphere_attach(R, 1e12, tuple(c), v=(V, 0.0, 0.0))
        for _ in range(150):
            m.step(H)
        for _ in range(6000):
            m.step(H)
            F.append(m.sphere_state()["force"][0])
            U.append(m.diagnostics()["momentum"][0] / len(pos))
        assert m.diagnostics()["flags"] == 0
    F, U = np.array(F), np.array(U)
    eta = 10 * OBS.srd_viscosity(10, H, mcm.CALPHA)
    x = R / L
    pred = -6 * np.pi * eta * R * (V - U.mean()) / (1 - 2.837297 * x + 4.19 * x ** 3 - 27.4 * x ** 6)
    blk = F.reshape(-1, 200).mean(1)
    ratio = F.mean() / pred
    print(f"[result] FP32-bulk drag {F.mean():.2f} +- {blk.std(ddof=1) / np.sqrt(len(blk)):.2f}, Stokes-Hasimoto "
          f"{pred:.2f}, ratio {ratio:.3f}", flush=True)
    assert 0.8 < ratio < 1.1
