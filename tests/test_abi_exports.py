"""The C-ABI library loads and exports every entry point include/mpcd_abi.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os

import mcm_amd


def test_library_exports_declared_symbols():
    assert os.path.exists(mcm_amd.LIB_PATH), "build the HIP library first"
    lib = ctypes.CDLL(mcm_amd.LIB_PATH)
    syms = mcm_amd.declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_error_string():
    L = mcm_amd.lib()
    assert L.mpcd_version().startswith(b"mpcd-gfx950")
    assert isinstance(L.mpcd_last_error(), bytes)


def test_bad_params_rejected_without_device():
    prm = mcm_amd.Params()
    prm.size[:] = [0, 4, 4]
    prm.rho = 10
    prm.kT = 1.0
    h = ctypes.c_void_p()
    assert mcm_amd.lib().mpcd_create(ctypes.byref(prm), ctypes.byref(h)) == -1
    assert b"box size" in mcm_amd.lib().mpcd_last_error()
