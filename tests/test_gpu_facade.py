"""The C++ facade with the reference's MPCD / Fluid method set
(include/gpu_mpcd.hpp) driven in Simulation::next_tstep's order with a
host-side MD object, bit for bit against the C ABI (tests/cpp/
facade_test.cpp, built by build())."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "multi-ciliated-microswimmers_amd", "lib", "facade_test")


def test_facade_executable_built():
    assert os.path.exists(EXE), "build() links lib/facade_test"


@pytest.mark.gpu
def test_facade_matches_c_abi(tmp_path):
    print("[case] facade_test", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "facade ok" in r.stdout
