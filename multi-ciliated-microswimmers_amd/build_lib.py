"""Builds the gfx950 MPCD library in-tree: lib/libmcm_mpcd.so, the
configuration-file driver lib/cilia and the facade's C++ test
lib/facade_test.

Each HIP translation unit is compiled with hipcc for gfx950 only, with
-ffp-contract=off (no implicit fused multiply-add: the FP64 form is
bit-comparable with the CPU restatement) and packed-FP32 instructions
disabled.  The driver's host sources (sim/) are plain C++ built with g++
into the same library; lib/cilia is its command-line front end.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libmcm_mpcd.so")
SIMDIR = os.path.join(HERE, "sim")
INCDIR = os.path.join(os.path.dirname(HERE), "include")
EXE = os.path.join(LIBDIR, "cilia")
# the C++ test of the MPCD / Fluid facade (include/gpu_mpcd.hpp), run by tests/test_gpu_facade.py
FACADE_TEST_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "facade_test.cpp")
FACADE_TEST = os.path.join(LIBDIR, "facade_test")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

SOURCES = ["mpcd_step_f32.hip", "mpcd_step_f64.hip", "mpcd_step_later_f32.hip", "mpcd_step_later_f64.hip",
           "mpcd_aux.hip", "mpcd_md.hip", "mpcd_abi.cpp"]
SIM_SOURCES = ["cfg.cpp", "objects.cpp", "simulation.cpp", "gpu_mpcd.cpp"]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCDIR}"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def _compile_host(src: str) -> str:
    obj = os.path.join(OBJDIR, "sim_" + os.path.splitext(src)[0] + ".o")
    _run([CXX, *CXXFLAGS, "-c", os.path.join(SIMDIR, src), "-o", obj], f"g++ failed on {src}")
    return obj


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(what)


def build() -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES) + len(SIM_SOURCES)) as ex:
        hip = [ex.submit(_compile, s) for s in SOURCES]
        host = [ex.submit(_compile_host, s) for s in SIM_SOURCES]
        objs = [f.result() for f in hip + host]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB, *objs], "link failed")
    _run([CXX, *CXXFLAGS, os.path.join(SIMDIR, "cilia_main.cpp"), "-o", EXE, f"-L{LIBDIR}", "-lmcm_mpcd",
          "-Wl,-rpath,$ORIGIN"], "link of cilia failed")
    _run([CXX, *CXXFLAGS, FACADE_TEST_SRC, "-o", FACADE_TEST, f"-L{LIBDIR}", "-lmcm_mpcd", "-Wl,-rpath,$ORIGIN"],
         "link of facade_test failed")
    return LIB


if __name__ == "__main__":
    print(build())
