"""Host-side check of the blend's sub-block pruning (splat_sub_block_mask in
torpedo_amd/csrc/tpd_device.hpp).

The blend drops a splat from the list of a 4x4 sub-block only when no pixel centre of that
sub-block can pass alpha >= 1/255 (blend.slang:22-105 tests `alpha < 1.0 / 255.0`).  The harness
tests/cpp/sub_block_mask.cpp compiles the device header for the host with hipcc and checks the
mask against brute force over all 256 pixels of the tile for random splats (projection-style and
raw, nearly degenerate conics; tiles near the origin and near 16384 px).  No GPU is used."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "sub_block_mask.cpp")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    hipcc = _hipcc()
    if hipcc is None:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("sbm") / "sub_block_mask")
    cmd = [hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
           "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "torpedo_amd", "csrc"), SRC, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("op_error", ["0", "-4.76837158203125e-07", "4.76837158203125e-07"])
def test_sub_block_mask_never_drops_a_reachable_pixel(harness, op_error):
    # op_error: relative error (0, -2^-21, +2^-21) given to the mask's square roots, reciprocals and
    # logarithm, which the device computes with approximate instructions (a few ulp)
    r = subprocess.run([harness, "400000", op_error], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    fields = dict(zip(r.stdout.split()[0::2], r.stdout.split()[1::2]))
    assert int(fields["violations"]) == 0
    # the splats that reach the tile at all are a sizeable share, so the check is not vacuous
    assert int(fields["reaching"]) > 40000
