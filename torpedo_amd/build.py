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
           "tuner.cpp", "k_util.hip", "noalloc.hip", "ply.cpp", "present.cpp", "capi.cpp"]


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


CPP_TESTS = ["engine_api", "hello_gaussian"]  # tests/cpp/<name>.cpp -> tests/cpp/<name>


def build_cpp_test() -> list[str]:
    """The C++ API test programs (include/torpedo/*.hpp over libtpd_hip.so), linked in-tree."""
    cxx = shutil.which("g++") or "g++"
    out = []
    for name in CPP_TESTS:
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        exe = os.path.join(ROOT, "tests", "cpp", name)
        cmd = [cxx, "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
               src, "-L", HERE, "-ltpd_hip", "-Wl,-rpath,$ORIGIN/../../torpedo_amd", "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed building tests/cpp/{name}.cpp")
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


def build_sort_compare() -> str:
    """tests/cpp/sort_vs_rocprim: the library's stand-alone sort against rocPRIM's radix sort."""
    src = os.path.join(ROOT, "tests", "cpp", "sort_vs_rocprim.hip")
    exe = os.path.join(ROOT, "tests", "cpp", "sort_vs_rocprim")
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-Wno-unused-result",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), src, "-L", HERE, "-ltpd_hip",
           "-Wl,-rpath,$ORIGIN/../../torpedo_amd", "-o", exe + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building tests/cpp/sort_vs_rocprim.hip")
    os.replace(exe + ".tmp", exe)
    return exe


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
