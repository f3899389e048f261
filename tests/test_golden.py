"""The oracle and the product's host code against golden vectors of the reference's own code.

tests/golden/*.npz hold inputs and outputs of the reference's camera/math code
(Camera.cpp:3-16 lookAt, PerspectiveCamera.cpp:3-23, the compensated mat4 multiply of mat4.h)
and of its PLY loader (GaussianGeometry.cpp:59-127), produced by tests/golden/make_golden.py from
the reference compiled in oracle/_ref.  These checks need neither the reference tree nor a GPU:
the camera functions and the PLY reader of libtpd_hip.so are host code."""
import ctypes
import os
import tempfile

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def cam():
    return np.load(os.path.join(GOLD, "camera_ref.npz"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_oracle_camera_matches_golden(orc, cam):
    for i in range(len(cam["eye"])):
        assert np.array_equal(_bits(orc.look_at(cam["eye"][i], cam["center"][i], cam["up"][i]).ravel()),
                              _bits(cam["look_at"][i])), i
        assert np.array_equal(_bits(orc.mat4_mul(cam["A"][i], cam["B"][i]).ravel()), _bits(cam["mat4_mul"][i])), i
    for i, (w, h) in enumerate(cam["sizes"]):
        for j, fov in enumerate(cam["fovs"]):
            fy = float(np.float32(np.sqrt(3.0))) if fov == 0.0 else orc.fy_from_fov(float(fov))
            assert np.array_equal(_bits(orc.perspective(int(w), int(h), fy).ravel()), _bits(cam["perspective"][i, j]))


def test_capi_camera_matches_golden(cam):
    from torpedo_amd._capi import lib
    L = lib()
    out = np.zeros(16, np.float32)
    for i in range(len(cam["eye"])):
        e, c, u = (np.ascontiguousarray(cam[k][i], np.float32) for k in ("eye", "center", "up"))
        L.tpd_look_at(_ptr(e), _ptr(c), _ptr(u), _ptr(out))
        assert np.array_equal(_bits(out), _bits(cam["look_at"][i])), i
        A, B = (np.ascontiguousarray(cam[k][i], np.float32) for k in ("A", "B"))
        L.tpd_mat4_mul(_ptr(A), _ptr(B), _ptr(out))
        assert np.array_equal(_bits(out), _bits(cam["mat4_mul"][i])), i
    for i, (w, h) in enumerate(cam["sizes"]):
        for j, fov in enumerate(cam["fovs"]):
            fy = float(np.float32(np.sqrt(3.0))) if fov == 0.0 else float(L.tpd_fy_from_fov(float(fov)))
            L.tpd_perspective(int(w), int(h), fy, 0.01, 100.0, _ptr(out))
            assert np.array_equal(_bits(out), _bits(cam["perspective"][i, j])), (w, h, fov)


def test_ply_reader_matches_golden():
    from torpedo_amd.engine import from_model
    z = np.load(os.path.join(GOLD, "ply_ref.npz"))
    names = sorted({k.split("/")[0] for k in z.files})
    assert len(names) >= 5
    with tempfile.TemporaryDirectory() as d:
        for name in names:
            path = os.path.join(d, name + ".ply")
            with open(path, "wb") as f:
                f.write(z[name + "/file"].tobytes())
            got = from_model(path)
            ref = z[name + "/records"]
            assert got.shape == ref.shape, name
            assert np.array_equal(_bits(got), _bits(ref)), name
