"""Builds libtpd_hip.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -m torpedo_amd.build          # compile
    python -m torpedo_amd.build --check  # print the command only

Flags that matter for results:
  -ffp-contract=off                             no implicit multiply-add fusion (bit parity with
                                                the CPU oracle)
  -fhip-fp32-correctly-rounded-divide-sqrt      IEEE division and square root
  -fno-gpu-flush-denormals-to-zero              IEEE subnormals
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtpd_hip.so")

SOURCES = ["k_forward.hip", "k_sort.hip", "k_backward.hip", "k_batch.hip", "render.cpp", "backward.cpp", "batch.cpp",
           "tuner.cpp", "k_util.hip", "ply.cpp", "present.cpp", "capi.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def command(out: str = LIB) -> list[str]:
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
            "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            # The kernels reduce across lanes themselves before every atomic to a wave-uniform
            # address; the compiler's own lane-by-lane reduction loop around such atomics only adds
            # instructions (one per active lane, per atomic).
            "-Xarch_device", "-mllvm=-amdgpu-atomic-optimizer-strategy=None",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC] + srcs + ["-o", out]


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "engine_api.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "engine_api")


def build_cpp_test() -> str:
    """The C++ API test program (include/torpedo/*.hpp over libtpd_hip.so), linked in-tree."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC,
           "-L", HERE, "-ltpd_hip", "-Wl,-rpath,$ORIGIN/../../torpedo_amd", "-o", CPP_TEST_BIN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the C++ API test")
    os.replace(CPP_TEST_BIN + ".tmp", CPP_TEST_BIN)
    return CPP_TEST_BIN


def build(verbose: bool = False) -> str:
    cmd = command(LIB + ".tmp")
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libtpd_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(" ".join(command()))
    else:
        print(build(verbose=True))
