"""The cell thermostat's output distribution on the device, both forms.

The MBS thermostat (mpcd.cpp:13-43, 200-212) rescales the velocities
relative to a cell's centre-of-mass velocity so that their kinetic energy
becomes kT G, with G drawn from Gamma(3(n-1)/2, 1) for a cell of n
particles; the rotation keeps that energy.  So after one step, for every
collision cell of n >= 2 particles,

    G = sum_i |v_i - u|^2 / (2 kT)

must be distributed as Gamma(3(n-1)/2).  Mapping each cell's G through the
Gamma CDF of its own shape gives values that must be uniform on (0, 1):
a Kolmogorov-Smirnov test over all cells checks the production draws end
to end (counter-based generator, Box-Muller normals, Marsaglia-Tsang
acceptance, the thermostat's scale factor), without injection.  The
positions before the step name each particle's collision cell (ids are
tracked, so the velocities after the step come back in the same order).
A planted error (the shape off by one) must fail the same test.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O

pytestmark = pytest.mark.gpu

KT = 1.0


def _cell_gamma(L, a, pos0, vel1):
    """Per collision cell (shift a) of n >= 2 particles: n and G."""
    cell = O.box(L, a, pos0)
    n = np.bincount(cell, minlength=int(np.prod(L)))
    sv = np.stack([np.bincount(cell, weights=vel1[:, k], minlength=len(n)) for k in range(3)], axis=1)
    u = sv / np.maximum(n, 1)[:, None]
    rel = np.bincount(cell, weights=((vel1[:, :3] - u[cell]) ** 2).sum(axis=1), minlength=len(n))
    keep = n >= 2
    return n[keep], rel[keep] / (2.0 * KT)


@pytest.mark.parametrize("form", ["fp32bulk", "fp64"])
def test_thermostat_gamma_distribution(mcm, form):
    L = (32, 32, 32)
    print(f"[case] thermostat Gamma KS form={form} L={L}", flush=True)
    rng = np.random.default_rng(11)
    n = 10 * int(np.prod(L))
    pos0 = rng.uniform(0.0, 1.0, (n, 3)) * np.array(L)
    vel0 = np.zeros((n, 4))
    vel0[:, :3] = rng.normal(0.0, 1.0, (n, 3))
    vel0[:, 3] = 1.0
    a0 = np.array([0.1, -0.3, 0.25])
    prec = mcm.FP32BULK if form == "fp32bulk" else mcm.FP64
    with mcm.MPCD(L, precision=prec, track_ids=True, seed=5) as m:
        m.set_shift(a0)
        m.upload_fluid(pos0, vel0)
        m.step(1e-6)
        _, vel1 = m.download_fluid()
    cnt, G = _cell_gamma(L, a0, pos0, vel1)
    u = stats.gamma.cdf(G, 1.5 * (cnt - 1))
    p = stats.kstest(u, "uniform").pvalue
    print(f"[result] {form}: {len(G)} cells, KS p = {p:.3g}", flush=True)
    assert p > 1e-3
    # a planted error: the shape of one fewer particle
    m3 = cnt >= 3
    p_bad = stats.kstest(stats.gamma.cdf(G[m3], 1.5 * (cnt[m3] - 2)), "uniform").pvalue
    assert p_bad < 1e-6
