"""The planar walls in the FP32-bulk form (the reference's SLIP / NOSLIP
build), on the GPU.

The FP32-bulk form takes the walls as the FP64 form does (the same virtual
mass in the bottom and top layers' cells below rho particles, from the same
draws; the same bounce-back, MPCD::move), storing a bounced particle's
velocity back on the 2^-21 grid.  The FP64 form is bit-exact against the
oracle fed the reference's MT19937 draws (test_gpu_walls.py), so the
FP32-bulk form is held to it:

* one step from the same state with the same injected draws (Gamma
  variates, axes, the virtual normals): positions and velocities within
  1e-5;
* a Couette flow between a resting bottom wall and a top wall moving at
  0.2: after the start-up, the time-averaged velocity profile of both forms
  agrees within its scatter and rises towards the moving wall, every
  particle stays between the walls, and the temperatures of the two forms
  agree.  Both forms show the reference's wall rules: its virtual
  particles carry no wall velocity, so the top layer's collisions drag the
  fluid back and the shear is about a tenth of U / (L_z - 1), and they run
  about 6% above kT.
"""
import numpy as np
import pytest

from cases import random_draws

pytestmark = pytest.mark.gpu

H = 0.1
Q = 2.0 ** 21


def _state(L, rng, rho=10):
    n = rho * L[0] * L[1] * (L[2] - 1)
    pos = rng.uniform(0, 1, (n, 3)) * np.array([L[0], L[1], L[2] - 1])
    vel = np.ones((n, 4))
    vel[:, :3] = np.rint(rng.normal(size=(n, 3)) * Q) / Q
    return pos, vel


def _min_image(d, L):
    L = np.asarray(L, dtype=float)
    return d - L * np.rint(d / L)


@pytest.mark.parametrize("walls,vw", [(1, (0.0, 0.0)), (1, (0.25, -0.1)), (2, (0.0, 0.0))])
def test_walls_step_fp32_against_fp64(mcm, walls, vw):
    L = (16, 12, 10)
    print(f"[case] walls_f32 walls={walls} vw={vw} L={L}", flush=True)
    rng = np.random.default_rng(100 + walls)
    pos, vel = _state(L, rng)
    ncells = int(np.prod(L))
    a = rng.uniform(-0.5, 0.5, 3)
    a_next = rng.uniform(-0.5, 0.5, 3)
    g, ax = random_draws(ncells, rng)
    virt = rng.normal(size=3 * ncells)
    out = {}
    for form, prec in (("fp64", mcm.FP64), ("fp32", mcm.FP32BULK)):
        with mcm.MPCD(L, precision=prec, track_ids=True, seed=7, walls=walls, wall_v=vw) as m:
            m.set_shift(a)
            m.upload_fluid(pos, vel)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax, inj_virt=virt)
            gp, gv = m.download_fluid()
            out[form] = (gp, gv, m.diagnostics()["flags"])
    p64, v64, f64 = out["fp64"]
    p32, v32, f32 = out["fp32"]
    assert f64 == 0 and f32 == 0
    dp = np.abs(_min_image(p32 - p64, L)).max()
    dv = np.abs(v32[:, :3] - v64[:, :3]).max()
    print(f"[result] walls={walls} vw={vw}: max |dx| {dp:.2e}, max |dv| {dv:.2e}", flush=True)
    assert dp < 1e-5 and dv < 1e-5
    assert p32[:, 2].min() >= 0.0 and p32[:, 2].max() <= L[2] - 1.0 + 1e-6


def test_couette_fp32_against_fp64(mcm):
    L = (16, 16, 17)
    U = 0.2
    print(f"[case] couette L={L} U={U}", flush=True)
    rng = np.random.default_rng(8)
    pos, vel = _state(L, rng)
    prof = {}
    for form, prec in (("fp64", mcm.FP64), ("fp32", mcm.FP32BULK)):
        with mcm.MPCD(L, precision=prec, seed=21, walls=1, wall_v=(U, 0.0)) as m:
            m.upload_fluid(pos, vel)
            for _ in range(3000):
                m.step(H)
            acc = []
            for _ in range(40):
                for _ in range(25):
                    m.step(H)
                p, v = m.download_fluid()
                assert p[:, 2].min() >= 0.0 and p[:, 2].max() <= L[2] - 1.0 + 1e-6
                z = np.clip(p[:, 2].astype(int), 0, L[2] - 2)
                acc.append(np.bincount(z, weights=v[:, 0], minlength=L[2] - 1) / np.bincount(z, minlength=L[2] - 1))
            d = m.diagnostics()
            assert d["flags"] == 0
            vv = v[:, :3] - v[:, :3].mean(0)
            kt = (vv ** 2).sum(1).mean() / 3.0
            prof[form] = (np.array(acc), kt)
    (a64, kt64), (a32, kt32) = prof["fp64"], prof["fp32"]
    m64, m32 = a64.mean(0), a32.mean(0)
    err = np.sqrt(a64.var(0, ddof=1) / len(a64) + a32.var(0, ddof=1) / len(a32))
    zc = np.arange(L[2] - 1) + 0.5
    slope = np.polyfit(zc[2:-2], m32[2:-2], 1)[0]
    print(f"[result] profile fp64 {np.round(m64, 3)}", flush=True)
    print(f"[result] profile fp32 {np.round(m32, 3)}", flush=True)
    print(f"[result] kT {kt64:.4f} / {kt32:.4f}, shear rate FP64 {np.polyfit(zc[2:-2], m64[2:-2], 1)[0]:.4f}, "
          f"FP32-bulk {slope:.4f} (U / (L_z - 1) = {U / (L[2] - 1):.4f})", flush=True)
    slope64 = np.polyfit(zc[2:-2], m64[2:-2], 1)[0]
    assert np.all(np.abs(m32 - m64) < 5 * err + 2e-3)
    assert abs(kt32 - kt64) < 0.01 and abs(kt32 - 1.0) < 0.08
    assert slope > 0.0 and abs(slope - slope64) < 0.3 * slope64 + 3e-4
