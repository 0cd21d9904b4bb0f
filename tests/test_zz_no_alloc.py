"""mpcd_create_fluid_noalloc (the creation that allocates nothing, its
scratch a fixed per-GPU array) against mpcd_create_fluid through the C ABI:
the same records bit for bit in both forms, on two calls in a row with the
second smaller (the shared scratch keeps the first call's partial sums),
the same state after a step, and a refusal of n beyond what the home
segments hold."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(mcm, m, n):
    L = mcm.lib()
    L.mpcd_create_fluid_noalloc.argtypes = [C.c_void_p, C.c_longlong]
    L.mpcd_create_fluid_noalloc.restype = C.c_int
    return L.mpcd_create_fluid_noalloc(m.h, int(n))


@pytest.mark.parametrize("prec_name", ["FP32BULK", "FP64"])
def test_noalloc_create_matches_create(mcm, prec_name):
    prec = getattr(mcm, prec_name)
    for size, n, seed in (((40, 36, 32), 10 * 40 * 36 * 32, 5), ((12, 10, 8), 777, 6)):
        print(f"[case] noalloc_create {prec_name} size={size} n={n}", flush=True)
        with mcm.MPCD(size, precision=prec, seed=seed, track_ids=True) as a, \
                mcm.MPCD(size, precision=prec, seed=seed, track_ids=True) as b:
            a.create_fluid(n)
            assert _noalloc(mcm, b, n) == 0, mcm.lib().mpcd_last_error()
            assert b.n == n
            pa, va = a.download_fluid()
            pb, vb = b.download_fluid()
            assert np.array_equal(pa, pb) and np.array_equal(va, vb)
            a.step(0.1)
            b.step(0.1)
            pa, va = a.download_fluid()
            pb, vb = b.download_fluid()
            assert np.array_equal(pa, pb) and np.array_equal(va, vb)
            assert a.diagnostics()["flags"] == 0 and b.diagnostics()["flags"] == 0


def test_noalloc_create_refuses_too_many(mcm):
    with mcm.MPCD((8, 8, 8), precision=mcm.FP32BULK, seed=1) as m:
        assert _noalloc(mcm, m, 10 ** 9) == -1
        assert b"home segments hold" in mcm.lib().mpcd_last_error()
