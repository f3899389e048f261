"""Writes the golden fixtures of tests/golden/ from the reference's own code.

The reference's camera/math code (torpedo/rendering/src/Camera.cpp, extension/src/
PerspectiveCamera.cpp, math/include/torpedo/math/*.h) and its PLY loader
(volumetric/src/GaussianGeometry.cpp:59-127 with its vendored miniply) are compiled from
/root/reference into oracle/_ref by oracle/ref_camera.sh and oracle/ref_ply.sh (run by
oracle.build()).  This script feeds them seeded inputs and stores inputs and outputs:

  camera_ref.npz   lookAt, compensated mat4 multiply and PerspectiveCamera projection
  ply_ref.npz      small 3DGS PLY files (every encoding the reader accepts) and the records
                   GaussianPoint::fromModel returns for them

so that machines without the reference tree (where oracle/_ref cannot be built) still check the
oracle and the product against the reference's own outputs.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import io
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from test_ply import columns, write_ply  # noqa: E402


def camera_fixture():
    ref = oracle.ref_camera()
    assert ref is not None, "oracle/_ref/libtpd_ref_camera.so is not built (run oracle.build())"
    rng = np.random.default_rng(20261202)
    n = 64
    eye = (rng.normal(size=(n, 3)) * 5).astype(np.float32)
    center = rng.normal(size=(n, 3)).astype(np.float32)
    up = rng.normal(size=(n, 3)).astype(np.float32)
    look = np.zeros((n, 16), np.float32)
    for i in range(n):
        ref.tpr_look_at(eye[i], center[i], up[i], look[i])
    A = rng.normal(size=(n, 16)).astype(np.float32)
    B = rng.normal(size=(n, 16)).astype(np.float32)
    mul = np.zeros((n, 16), np.float32)
    for i in range(n):
        ref.tpr_mat4_mul(A[i], B[i], mul[i])
    sizes = np.array([(1280, 720), (1920, 1080), (256, 256), (1600, 900), (333, 77)], np.uint32)
    fovs = np.array([0.0, 45.0, 60.0, 90.0], np.float32)
    persp = np.zeros((len(sizes), len(fovs), 16), np.float32)
    for i, (w, h) in enumerate(sizes):
        for j, f in enumerate(fovs):
            ref.tpr_perspective(int(w), int(h), float(f), persp[i, j])
    np.savez_compressed(os.path.join(HERE, "camera_ref.npz"), eye=eye, center=center, up=up, look_at=look, A=A, B=B,
                        mat4_mul=mul, sizes=sizes, fovs=fovs, perspective=persp)


CASES = [  # name, n, seed, sh rest coefficients, format, property type, leading non-vertex element
    ("le_deg3", 7, 1, 45, "binary_little_endian", "float", 0),
    ("be_deg3", 5, 2, 45, "binary_big_endian", "float", 0),
    ("ascii_deg1", 4, 3, 9, "ascii", "float", 0),
    ("le_deg0_double", 6, 4, 0, "binary_little_endian", "double", 0),
    ("le_deg2", 3, 5, 24, "binary_little_endian", "float", 0),
]


def ply_fixture():
    assert oracle.ref_ply() is not None, "oracle/_ref/libtpd_ref_ply.so is not built (run oracle.build())"
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, n, seed, n_rest, fmt, dt, faces in CASES:
            path = os.path.join(d, name + ".ply")
            write_ply(path, columns(n, seed, n_rest), fmt=fmt, pre_faces=faces, dtype=dt)
            rec = oracle.ref_from_model(path)
            assert rec is not None and rec.shape == (n, 60), name
            out[name + "/file"] = np.frombuffer(open(path, "rb").read(), np.uint8)
            out[name + "/records"] = rec
    np.savez_compressed(os.path.join(HERE, "ply_ref.npz"), **out)


if __name__ == "__main__":
    oracle.build()
    camera_fixture()
    ply_fixture()
    for f in ("camera_ref.npz", "ply_ref.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
