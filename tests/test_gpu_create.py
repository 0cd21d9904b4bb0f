"""Device fluid creation (Fluid::create, src/libcilia/fluid.cpp:55-98,
mpcd_create_fluid) against the oracle's restatement of the same Philox
'FLUD' stream, and against the distributions the reference draws from:

* FP64: positions bit-identical to the oracle's (one rounded product per
  coordinate on both sides); velocities within 1e-12 (the device's and
  glibc's log/cos/sin differ by an ulp or so) after the mean is removed.
* FP32-bulk: the same state through the cell-offset encoding (positions
  within 2^-20 of a cell, velocities on the 2^-21 grid).
* Kolmogorov-Smirnov: positions uniform over [0, L) on each axis,
  velocities N(0, kT) per component at kT = 1 and 2; mean velocity zero.
* Planar walls: every particle is kept, z over [0, Lz) as the reference's
  SLIP / NOSLIP build draws it, and one step puts all of them between the
  walls.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O

pytestmark = pytest.mark.gpu


def _print_case(**kw):
    print("[case] create", kw, flush=True)


@pytest.mark.parametrize("prec_name", ["FP64", "FP32BULK"])
def test_create_matches_oracle_stream(mcm, prec_name):
    L, seed, kT = (24, 20, 16), 77, 1.0
    n = 10 * int(np.prod(L))
    _print_case(prec=prec_name, L=L, seed=seed)
    prec = getattr(mcm, prec_name)
    rp, rv = O.fluid_create_philox(seed, L, n, kT)
    with mcm.MPCD(L, precision=prec, seed=seed, kT=kT, track_ids=True) as m:
        m.create_fluid(n)
        assert m.n == n
        pos, vel = m.download_fluid()
    assert np.all(vel[:, 3] == 1.0)
    if prec_name == "FP64":
        assert np.array_equal(pos, rp)
        assert np.abs(vel[:, :3] - rv[:, :3]).max() < 1e-12
    else:
        d = pos - rp
        d -= np.asarray(L) * np.round(d / np.asarray(L))
        assert np.abs(d).max() < 2.0 ** -20
        assert np.abs(vel[:, :3] - rv[:, :3]).max() <= 2.0 ** -21


@pytest.mark.parametrize("kT", [1.0, 2.0])
def test_create_distributions(mcm, kT):
    L, seed = (32, 32, 32), 5
    n = 10 * 32 ** 3
    _print_case(L=L, seed=seed, kT=kT)
    with mcm.MPCD(L, precision=mcm.FP64, seed=seed, kT=kT) as m:
        m.create_fluid(n)
        pos, vel = m.download_fluid()
    for a in range(3):
        assert stats.kstest(pos[:, a] / L[a], "uniform").pvalue > 1e-3
        assert stats.kstest(vel[:, a] / np.sqrt(kT), "norm").pvalue > 1e-3
    assert np.abs(vel[:, :3].mean(axis=0)).max() < 1e-13
    assert np.mean(vel[:, :3] ** 2) == pytest.approx(kT, rel=5e-3)


def test_create_with_walls(mcm):
    L, seed = (16, 16, 8), 3
    n = 10 * int(np.prod(L))
    _print_case(L=L, seed=seed, walls=1)
    with mcm.MPCD(L, precision=mcm.FP64, seed=seed, walls=1) as m:
        m.create_fluid(n)
        assert m.n == n
        pos, _ = m.download_fluid()
        assert pos[:, 2].min() >= 0.0 and pos[:, 2].max() < L[2]
        above = np.mean(pos[:, 2] > L[2] - 1)
        assert above == pytest.approx(1.0 / L[2], abs=0.01)
        m.step(0.1)
        pos, _ = m.download_fluid()
        assert m.n == n
        assert pos[:, 2].min() >= 0.0 and pos[:, 2].max() <= L[2] - 1
