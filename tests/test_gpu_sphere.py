"""GPU parity of the rigid sphere (FP64 form; not in the reference; the
FP32-bulk form is held to this one in test_gpu_sphere_f32.py).

With the per-cell Gamma variates and axes and the sphere's virtual-particle
normals injected, one step's solvent positions and velocities must be
bit-identical to the oracle's (collision with virtual particles, stream
with bounce-back); the sphere's new velocities sum per-particle impulses in
another order, so they agree to 1e-12.  Over many steps on a larger box:
solvent + sphere momentum is conserved to roundoff and no solvent particle
ends inside the sphere.
"""
import numpy as np
import pytest

import oracle as O
from cases import random_draws

pytestmark = pytest.mark.gpu

H = 0.1


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _fluid_outside(L, X, R, rng, rho=10):
    n = rho * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    d = pos - np.asarray(X)
    d -= np.array(L) * np.rint(d / np.array(L))
    pos = pos[(d ** 2).sum(1) > R * R]
    vel = np.ones((len(pos), 4))
    vel[:, :3] = rng.normal(size=(len(pos), 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    return pos, vel


def _inside(pos, X, R, L):
    d = pos - np.asarray(X)
    d -= np.array(L) * np.rint(d / np.array(L))
    return (d ** 2).sum(1) < R * R


@pytest.mark.parametrize("slip", [False, True])
def test_sphere_step_bit_exact(mcm, slip):
    L = (16, 12, 12)
    X, R = (8.3, 5.7, 6.2), 3.1
    V, W = (0.05, -0.03, 0.02), (0.1, -0.2, 0.15)
    Ms = 4.0 / 3.0 * np.pi * R ** 3 * 10
    seed = 12
    _print_case("sphere_step", slip=slip, L=L)
    rng = np.random.default_rng(41)
    pos, vel = _fluid_outside(L, X, R, rng)
    ncells = int(np.prod(L))
    sp = O.make_sphere(R, Ms, X, V=V, W=W, slip=slip)
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=seed) as m:
        a = rng.uniform(0, 1, 3)
        m.set_shift(a)
        m.upload_fluid(pos, vel)
        m.sphere_attach(R, Ms, X, V, W, slip=slip)
        for s in range(3):
            g, ax = random_draws(ncells, rng)
            sph = rng.normal(size=3 * 10 * ncells)
            sp.inj_sph = sph.ctypes.data_as(O.C.POINTER(O.C.c_double))
            a_next = rng.uniform(0, 1, 3)
            cfg, keep, _ = O.make_cfg(L, rng=O.RNG_INJECT, gamma=g, axis=ax, seed=seed, step=s, sphere=sp)
            O.collide(cfg, a, pos, vel)
            O.stream_sphere(pos, vel, H, L, sp)
            O.sphere_finish(sp, H, L)
            m.step(H, next_shift=a_next, inj_gamma=g, inj_axis=ax, inj_sph=sph)
            gp, gv = m.download_fluid()
            st = m.sphere_state()
            if s == 0:
                assert np.array_equal(gp, pos), f"positions differ, max {np.abs(gp - pos).max()}"
                assert np.array_equal(gv[:, :3], vel[:, :3])
            else:  # the sphere's state now differs in the last bits
                assert np.abs(gp - pos).max() < 1e-9
                assert np.abs(gv[:, :3] - vel[:, :3]).max() < 1e-9
            np.testing.assert_allclose(st["v"], sp.V[:], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(st["w"], sp.W[:], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(st["x"], sp.X[:], rtol=1e-12, atol=1e-12)
            assert m.diagnostics()["flags"] == 0
            a = a_next


def test_sphere_large_box_invariants(mcm):
    L = (32, 32, 32)
    X, R = (16.0, 15.5, 16.5), 5.0
    Ms = 4.0 / 3.0 * np.pi * R ** 3 * 10
    _print_case("sphere_large", L=L)
    rng = np.random.default_rng(7)
    pos, vel = _fluid_outside(L, X, R, rng)
    with mcm.MPCD(L, precision=mcm.FP64, seed=3) as m:
        m.upload_fluid(pos, vel)
        m.sphere_attach(R, Ms, X, v=(0.2, 0.0, 0.0))
        P0 = vel[:, :3].sum(0) + Ms * np.array([0.2, 0.0, 0.0])
        for _ in range(50):
            m.step(H)
        st = m.sphere_state()
        d = m.diagnostics()
        assert d["flags"] == 0
        P1 = d["momentum"] + Ms * st["v"]
        np.testing.assert_allclose(P1, P0, rtol=0, atol=1e-8)
        gp, _ = m.download_fluid()
        assert _inside(gp, st["x"], R, L).sum() == 0
        assert st["v"][0] < 0.2  # the solvent drags the sphere


def test_sphere_refused_where_unsupported(mcm):
    with mcm.MPCD((16, 16, 16), precision=mcm.FP64, walls=mcm.NOSLIP) as m:
        with pytest.raises(mcm.MPCDError):
            m.sphere_attach(3.0, 100.0, (8.0, 8.0, 8.0))
    with mcm.MPCD((16, 16, 16), precision=mcm.FP64) as m:
        with pytest.raises(mcm.MPCDError):
            m.sphere_attach(8.0, 100.0, (8.0, 8.0, 8.0))  # too large for the box


def test_stokes_drag(mcm):
    """A rigid no-slip sphere (radius 4) moved at a fixed velocity V = 0.1
    (mass 1e12) through the fluid of a periodic 48^3 box: the mean force
    over 6000 steps after 150 matches Stokes' drag 6 pi eta R (V - U) with
    Hasimoto's periodic-array correction 1 / (1 - 2.8373 x + 4.19 x^3 -
    27.4 x^6), x = R / L (J. Fluid Mech. 5 (1959) 317), eta = rho nu the SRD
    viscosity of the closed form, U the fluid's mean velocity: ratio
    0.93 +- 0.04 measured; within 0.8 .. 1.1."""
    from mcm_amd import observables as OBS
    L, R, V = 48, 4.0, 0.1
    _print_case("stokes_drag", L=L, R=R)
    rng = np.random.default_rng(5)
    n = 10 * L ** 3
    pos = rng.uniform(0, 1, (n, 3)) * L
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    vel[:, :3] -= vel[:, :3].mean(0)
    c = np.array([L / 2.0] * 3)
    keep = ((pos - c) ** 2).sum(1) > R * R
    pos, vel = pos[keep], vel[keep]
    F, U = [], []
    with mcm.MPCD((L, L, L), precision=mcm.FP64, seed=3) as m:
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
    print(f"[result] drag {F.mean():.2f} +- {blk.std(ddof=1) / np.sqrt(len(blk)):.2f}, Stokes-Hasimoto {pred:.2f}, "
          f"ratio {ratio:.3f}", flush=True)
    assert 0.8 < ratio < 1.1
