"""Checkpoint and restart on the GPU.

A run that checkpoints after k steps and continues must be bit-identical,
step for step, to a fresh context (same parameters, the same swimmer or
sphere attached in any state) that loads the checkpoint and continues:
solvent positions and velocities by particle id, the swimmer's positions,
velocities, rest lengths, beat state and globalt; the running dissipation
total (summed with atomics) to 1e-12.  This pins the stored state as
complete: the random streams (keyed by seed and step), the shift, the
records (exact in both forms, also where an FP32 offset is tiny) and the MD
state.  A rigid sphere restarts as described in its test.
"""
import numpy as np
import pytest

from mcm_amd import swimmer as SW

pytestmark = pytest.mark.gpu

H = 0.1
DT = 0.002


def _print_case(name, **kw):
    print(f"[case] {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _fluid(L, rng):
    n = 10 * int(np.prod(L))
    pos = rng.uniform(0, 1, (n, 3)) * np.array(L)
    pos[:5, 0] = 3.0 + np.array([1e-9, 3e-8, 2e-7, 0.0, 5e-12])  # tiny FP32 offsets (exact round trip)
    vel = np.ones((n, 4))
    vel[:, :3] = rng.normal(size=(n, 3))
    return pos, vel


def _state(m, sw):
    p, v = m.download_fluid()
    d = m.diagnostics()
    out = {"p": p, "v": v, "diss": d["diss_total"]}
    if sw:
        md = m.md_fetch()
        out.update({"bp": md["pos"], "bv": md["vel"], "nl": md["newlength"], "ol": md["oldlength"],
                    "beat": md["beat"], "gt": md["globalt"]})
    return out


def _same(a, b):
    for k in a:
        if k == "diss":  # summed with FP64 atomics: the order, and the last bit, vary
            assert abs(a[k] - b[k]) <= 1e-12 * abs(a[k]), "dissipation total differs after the restart"
        elif isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), f"{k} differs after the restart"
        else:
            assert a[k] == b[k], f"{k} differs after the restart"


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_restart_continues_bit_for_bit(mcm, prec_name, tmp_path):
    L = (20, 16, 24)
    prec = getattr(mcm, prec_name)
    _print_case("restart", prec=prec_name, L=L)
    rng = np.random.default_rng(17)
    pos, vel = _fluid(L, rng)
    desc = SW.clami(levels=3, radius=4.0, center=(10.0, 8.0, 12.0), n_cilia=8, rodlength=4, seed=2)
    path = str(tmp_path / "run.ckpt")
    with mcm.MPCD(L, precision=prec, track_ids=True, seed=31) as a:
        a.upload_fluid(pos, vel)
        a.md_attach(desc, globalt=0.0, ststep=DT)
        for _ in range(4):
            a.step(H)
        a.checkpoint_save(path)
        after = []
        for _ in range(3):
            a.step(H)
            after.append(_state(a, True))
    with mcm.MPCD(L, precision=prec, track_ids=True, seed=99) as b:  # seed comes from the file
        b.create_fluid(100)  # any prior state
        b.md_attach(desc, globalt=5.0, ststep=DT)
        b.checkpoint_load(path)
        for s in range(3):
            b.step(H)
            _same(after[s], _state(b, True))


def test_restart_with_sphere(mcm, tmp_path):
    """The sphere's impulses are summed with atomics, so runs with a sphere
    agree to roundoff, not bit for bit, once its state has moved (as in the
    sphere's own parity test): the first step after the restart starts from
    identical states and is bit-identical in the solvent; the next within
    1e-9, the sphere within 1e-12."""
    L = (16, 16, 16)
    _print_case("restart_sphere", L=L)
    rng = np.random.default_rng(18)
    pos, vel = _fluid(L, rng)
    d = pos - np.array([8.0, 8.0, 8.0])
    keep = (d ** 2).sum(1) > 3.0 ** 2
    pos, vel = pos[keep], vel[keep]
    Ms = 4.0 / 3.0 * np.pi * 27.0 * 10
    path = str(tmp_path / "sph.ckpt")
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=4) as a:
        a.upload_fluid(pos, vel)
        a.sphere_attach(3.0, Ms, (8.0, 8.0, 8.0), v=(0.1, 0.0, 0.0))
        for _ in range(3):
            a.step(H)
        a.checkpoint_save(path)
        refs = []
        for _ in range(2):
            a.step(H)
            refs.append((a.download_fluid(), a.sphere_state()))
    with mcm.MPCD(L, precision=mcm.FP64, track_ids=True, seed=4) as b:
        b.create_fluid(10)
        b.sphere_attach(3.0, Ms, (1.0, 2.0, 3.0))
        b.checkpoint_load(path)
        for s, ((rp, rv), rs) in enumerate(refs):
            b.step(H)
            gp, gv = b.download_fluid()
            st = b.sphere_state()
            if s == 0:
                assert np.array_equal(gp, rp) and np.array_equal(gv, rv)
            else:
                assert np.abs(gp - rp).max() < 1e-9 and np.abs(gv - rv).max() < 1e-9
            for k in ("x", "v", "w"):
                np.testing.assert_allclose(st[k], rs[k], rtol=1e-12, atol=1e-14)


def test_checkpoint_refusals(mcm, tmp_path):
    _print_case("checkpoint_refusals")
    path = str(tmp_path / "x.ckpt")
    with mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK, seed=1) as a:
        a.create_fluid(10 * 16 ** 3)
        a.step(H)
        a.checkpoint_save(path)
    with mcm.MPCD((16, 16, 20), precision=mcm.FP32BULK) as b:
        with pytest.raises(mcm.MPCDError):
            b.checkpoint_load(path)  # another box
    with mcm.MPCD((16, 16, 16), precision=mcm.FP64) as b:
        with pytest.raises(mcm.MPCDError):
            b.checkpoint_load(path)  # another form
    with mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK) as b:
        b.md_attach(SW.clami(levels=3, radius=3.0, center=(8.0, 8.0, 8.0), n_cilia=4, rodlength=4), ststep=DT)
        with pytest.raises(mcm.MPCDError):
            b.checkpoint_load(path)  # a swimmer the checkpoint does not have
    with mcm.MPCD((16, 16, 16), precision=mcm.FP32BULK) as b:
        with pytest.raises(mcm.MPCDError):
            b.checkpoint_load(str(tmp_path / "missing.ckpt"))
