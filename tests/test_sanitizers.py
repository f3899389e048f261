"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, no GPU).

  PLY reader    torpedo_amd/csrc/ply.cpp reads untrusted files: built host-only with clang's
                sanitizers into tests/cpp/ply_sanitize.cpp and run over valid files in every
                encoding and over corrupted ones (truncated at many points, bytes flipped in the
                header and the data, absurd element and list counts, missing header lines); any
                sanitizer report aborts the run;
  blend mask    tests/cpp/sub_block_mask.cpp (the device header's sub-block mask compiled for the
                host) on 20,000 random splats;
  scene order   tests/cpp/scene_order.cpp (the C++ Scene mirror) built with g++'s sanitizers.
Every build here compiles host code only (clang's --cuda-host-only, or g++): no GPU code is built
with a sanitizer.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_ply import columns, write_ply

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")


def host_hip_build(tmp_path, name, sources, extra=()):
    if not os.path.exists(CLANG):
        pytest.skip("ROCm clang++ not found")
    exe = str(tmp_path / name)
    cmd = [CLANG, "-x", "hip", "--cuda-host-only", "--offload-arch=gfx950", "--rocm-path=/opt/rocm", "-std=c++17",
           *SAN, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "torpedo_amd", "csrc"), *extra,
           *sources, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def corrupt_variants(rng, data, header_len):
    out = []
    for cut in sorted(set(np.linspace(1, len(data) - 1, 40).astype(int).tolist() + [header_len - 1, header_len,
                                                                                      header_len + 1])):
        out.append(data[:cut])
    for _ in range(60):  # flipped bytes in the header
        b = bytearray(data)
        for _ in range(rng.integers(1, 4)):
            b[rng.integers(0, header_len)] = int(rng.integers(0, 256))
        out.append(bytes(b))
    for _ in range(40):  # flipped bytes in the data
        b = bytearray(data)
        for _ in range(rng.integers(1, 16)):
            b[rng.integers(header_len, len(b))] = int(rng.integers(0, 256))
        out.append(bytes(b))
    return out


def test_ply_reader_under_sanitizers(tmp_path):
    exe = host_hip_build(tmp_path, "ply_sanitize",
                         [os.path.join(ROOT, "torpedo_amd", "csrc", "ply.cpp"),
                          os.path.join(ROOT, "tests", "cpp", "ply_sanitize.cpp")])
    rng = np.random.default_rng(17)
    files, expect = [], []
    c = columns(25, seed=3)
    for fmt in ("binary_little_endian", "binary_big_endian", "ascii"):
        for faces, dtype in ((0, "float"), (3, "float"), (0, "double")):
            p = tmp_path / f"ok_{fmt}_{faces}_{dtype}.ply"
            write_ply(p, c, fmt=fmt, pre_faces=faces, dtype=dtype)
            files.append(str(p))
            expect.append(25)
            data = open(p, "rb").read()
            hl = data.index(b"end_header\n") + len(b"end_header\n")
            for k, v in enumerate(corrupt_variants(rng, data, hl)):
                q = tmp_path / f"bad_{fmt}_{faces}_{dtype}_{k}.ply"
                open(q, "wb").write(v)
                files.append(str(q))
                expect.append(None)
    head = ("ply\nformat binary_little_endian 1.0\nelement face 4000000000\nproperty list uint int idx\n"
            "element vertex 1\n" + "".join(f"property float {k}\n" for k in c) + "end_header\n").encode()
    p = tmp_path / "huge_list.ply"
    open(p, "wb").write(head + b"\xff" * 64)
    files += [str(p), str(tmp_path / "missing.ply")]
    expect += [None, None]
    r = subprocess.run([exe] + files, capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr[-6000:]
    lines = r.stdout.strip().split("\n")
    assert len(lines) == len(files)
    bad_ok = 0
    for f, want, line in zip(files, expect, lines):
        rc, n = int(line.split()[0]), int(line.split()[1])
        if want is not None:
            assert rc == 0 and n == want, (f, line)
        else:
            bad_ok += rc == 0
    print(f"{len(files)} files, {bad_ok} corrupted ones still parsed as PLY")


def test_sub_block_mask_under_sanitizers(tmp_path):
    exe = host_hip_build(tmp_path, "sub_block_mask", [os.path.join(ROOT, "tests", "cpp", "sub_block_mask.cpp")],
                         extra=["-ffp-contract=off"])
    r = subprocess.run([exe, "20000", "0"], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    fields = dict(zip(r.stdout.split()[0::2], r.stdout.split()[1::2]))
    assert int(fields["violations"]) == 0


def test_scene_order_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "scene_order")
    cmd = [gxx, "-std=c++20", *SAN, "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "scene_order.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.split("\n")[0].split() == ["20", "21", "22", "10", "11", "8", "7"]
