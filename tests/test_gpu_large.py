"""Size-independent properties of the step at large sizes on the GPU.

At sizes the oracle cannot follow in seconds, the step is checked through
what the physics fixes: particle count, exact momentum (FP32-bulk) or
momentum to roundoff (FP64), the thermostat's temperature, and clean
capacity flags after many steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_large_box_invariants(mcm, prec_name):
    L = (128, 128, 128)
    prec = getattr(mcm, prec_name)
    print(f"[case] large_box {prec_name} L={L}", flush=True)
    n = 10 * int(np.prod(L))
    with mcm.MPCD(L, precision=prec, seed=2024) as m:
        m.create_fluid(n)
        d0 = m.diagnostics()
        for _ in range(20):
            m.step(0.1)
        d = m.diagnostics()
        assert d["flags"] == 0
        assert d["n"] == n
        kT = d["twice_kinetic"] / (3.0 * n)
        assert abs(kT - 1.0) < 5e-3
        if prec == mcm.FP32BULK:
            assert np.all(np.abs(d["momentum_q"]) < 2 ** 40)
            m.step(0.1)
            d2 = m.diagnostics()
            assert np.array_equal(d2["momentum_q"], d["momentum_q"])
        else:
            assert np.max(np.abs(d["momentum"])) < 1e-6 * n ** 0.5
