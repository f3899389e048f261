"""The C ABI library loads, exports every symbol include/tpd_hip.h declares, and its host-side
camera math is bit-identical to the oracle's restatement (and, where built, to the reference's
own camera code).  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    src = open(os.path.join(ROOT, "include", "tpd_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tpd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from torpedo_amd import _capi
    lib = _capi.lib()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in _capi.SIGNATURES, f"{name} has no ctypes signature"


def test_host_camera_math_matches_oracle_and_reference(orc):
    from torpedo_amd import _capi
    lib = _capi.lib()
    ref = orc.ref_camera()
    rng = np.random.default_rng(1)
    for _ in range(100):
        eye, cen, up = (rng.normal(size=3).astype(np.float32) * s for s in (6, 1, 1))
        v = np.zeros(16, np.float32)
        lib.tpd_look_at(_capi.ptr(eye), _capi.ptr(cen), _capi.ptr(up), _capi.ptr(v))
        assert np.array_equal(v.view(np.uint32), orc.look_at(eye, cen, up).ravel().view(np.uint32))
        if ref is not None:
            r = np.zeros(16, np.float32)
            ref.tpr_look_at(eye, cen, up, r)
            assert np.array_equal(v.view(np.uint32), r.view(np.uint32))
        A, B = rng.normal(size=16).astype(np.float32), rng.normal(size=16).astype(np.float32)
        m = np.zeros(16, np.float32)
        lib.tpd_mat4_mul(_capi.ptr(A), _capi.ptr(B), _capi.ptr(m))
        assert np.array_equal(m.view(np.uint32), orc.mat4_mul(A, B).ravel().view(np.uint32))
    for w, h, fov in [(1280, 720, 60.0), (1920, 1080, 45.0), (512, 512, 90.0)]:
        fy = lib.tpd_fy_from_fov(fov)
        assert fy == np.float32(orc.fy_from_fov(fov))
        p = np.zeros(16, np.float32)
        lib.tpd_perspective(w, h, fy, 0.01, 100.0, _capi.ptr(p))
        assert np.array_equal(p, orc.perspective(w, h, fy).ravel())
        view = orc.look_at([1, 2, -3], [0, 0, 0], [0, -1, 0]).ravel()
        pv = np.zeros(16, np.float32)
        fo = np.zeros(2, np.float32)
        lib.tpd_camera_upload(_capi.ptr(view), _capi.ptr(p), _capi.ptr(pv), _capi.ptr(fo))
        opv, ofo = orc.camera_upload(view, p)
        assert np.array_equal(pv, opv.ravel()) and np.array_equal(fo, ofo)


def test_python_camera_mirrors_reference_api(orc):
    from torpedo_amd.engine import PerspectiveCamera
    cam = PerspectiveCamera(1920, 1080)
    cam.lookAt([0, 0, -8], [0, 0, 0], [0, -1, 0])
    view, pv, fo = cam.uniforms()
    o = orc.make_camera([0, 0, -8], [0, 0, 0], [0, -1, 0], 1920, 1080)
    assert np.array_equal(view, o.view) and np.array_equal(pv, o.proj_view) and np.array_equal(fo, o.focal_ndc)
    cam.setVerticalFov(45.0)
    assert np.float32(cam.getProjection()[1, 1]) == np.float32(orc.fy_from_fov(45.0))


def test_scene_orders_groups_before_singles():
    from torpedo_amd.engine import Scene, gaussian_point
    s = Scene()
    single = gaussian_point([9, 9, 9], 1.0)
    e0 = s.add(single)
    grp = np.zeros((3, 60), np.float32)
    grp[:, 0] = [1, 2, 3]
    e1 = s.add(grp)
    assert s.countAll() == 4 and s.count() == 1 and s.countGroup() == 1
    data = s.dataAll()
    assert list(data[:, 0]) == [1, 2, 3, 9]
    assert s.buildEntityMap() == {e1: 0, e0: 1}
    assert list(s.entityIndices()) == [0, 0, 0, 1]


def test_scene_lists_each_kind_last_added_first():
    """Scene.h:130-162 walk single-storage EnTT views, which visit the storage's packed array from
    its last element to its first: within groups and within singles, the last added comes first,
    and dataAll, groupSizes, buildEntityMap and the per-point entity indices agree on that order."""
    from torpedo_amd.engine import Scene, gaussian_point
    s = Scene()
    g1 = np.zeros((2, 60), np.float32)
    g1[:, 0] = [10, 11]
    g2 = np.zeros((3, 60), np.float32)
    g2[:, 0] = [20, 21, 22]
    ea = s.add(g1)
    es1 = s.add(gaussian_point([7, 0, 0], 1.0))
    eb = s.add(g2)
    es2 = s.add(gaussian_point([8, 0, 0], 1.0))
    assert (ea, es1, eb, es2) == (0, 1, 2, 3)
    assert list(s.dataAll()[:, 0]) == [20, 21, 22, 10, 11, 8, 7]
    assert s.groupSizes() == [3, 2]
    assert s.buildEntityMap() == {eb: 0, ea: 1, es2: 2, es1: 3}
    assert list(s.entityIndices()) == [0, 0, 0, 1, 1, 2, 3]


def test_create_without_gpu_reports_no_device():
    from torpedo_amd import _capi
    lib = _capi.lib()
    h = ctypes.c_void_p()
    cfg = _capi.TpdConfig()
    cfg.device = 0
    rc = lib.tpd_create(ctypes.byref(h), ctypes.byref(cfg))
    if rc == 0:
        lib.tpd_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == 7 and b"no HIP device" in lib.tpd_last_error()
    cfg.device = -1
    assert lib.tpd_create(ctypes.byref(h), ctypes.byref(cfg)) in (1, 7)
    assert lib.tpd_render(None, None) == 1


def test_cpp_scene_lists_each_kind_last_added_first(tmp_path):
    """The same order from the C++ mirror (include/torpedo/Scene.hpp), built host-only with g++."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path / "scene_order")
    cmd = [gxx, "-std=c++20", "-O1", "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           os.path.join(ROOT, "tests", "cpp", "scene_order.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60, check=True).stdout.split("\n")
    # entities in creation order: group (10, 11), single 7, group (20, 21, 22), single 8
    assert out[0].split() == ["20", "21", "22", "10", "11", "8", "7"]
    assert out[1].split() == ["3", "2"]
    assert out[2].split() == ["1", "3", "0", "2"]
    assert out[3].split() == ["0", "0", "0", "1", "1", "2", "3"]
