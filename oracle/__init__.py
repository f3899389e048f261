"""CPU oracle for torpedo's Gaussian rasterization path (ctypes wrapper over libtpd_oracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this package, and only as the checker.  The product library in
``torpedo_amd/`` never imports it.

The C sources (tpd_oracle.c, oracle_core.inc) restate the reference's Slang kernels and host
camera math; see the header of tpd_oracle.c for the file:line map.  ``ref_camera()`` returns the
reference's own camera/math code compiled by ref_camera.sh (oracle/_ref), when it was built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def build() -> None:
    """Compile libtpd_oracle.so (and the reference camera library when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/torpedo"):
        subprocess.run([os.path.join(_HERE, "ref_camera.sh")], check=True, stdout=subprocess.DEVNULL)
        subprocess.run([os.path.join(_HERE, "ref_ply.sh")], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libtpd_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.tpo_expf.restype = ctypes.c_float
        L.tpo_expf.argtypes = [ctypes.c_float]
        L.tpo_expf_vec.argtypes = [_f32p, _f32p, ctypes.c_uint64]
        L.tpo_expf_sweep.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
        L.tpo_look_at.argtypes = [_f32p, _f32p, _f32p, _f32p]
        L.tpo_perspective.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, _f32p]
        L.tpo_fy_from_fov.restype = ctypes.c_float
        L.tpo_fy_from_fov.argtypes = [ctypes.c_float]
        L.tpo_mat4_mul.argtypes = [_f32p, _f32p, _f32p]
        L.tpo_camera_upload.argtypes = [_f32p, _f32p, _f32p, _f32p]
        L.tpo_preprocess_f32.argtypes = [_f32p, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _f32p, _f32p, _f32p,
                                         ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _f32p,
                                         _u32p, _i32p, ctypes.c_int]
        L.tpo_scan.restype = ctypes.c_uint64
        L.tpo_scan.argtypes = [_u32p, ctypes.c_uint32, _u32p]
        L.tpo_duplicate.argtypes = [_f32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _u64p, _u32p,
                                    ctypes.c_int]
        L.tpo_sort_pairs.argtypes = [_u64p, _u32p, ctypes.c_uint64, ctypes.c_int]
        L.tpo_ranges.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint32, _u32p]
        L.tpo_blend_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32,
                                    ctypes.c_uint32, _vp, _vp, ctypes.c_int]
        L.tpo_blend_decisions_f32.restype = ctypes.c_uint64
        L.tpo_blend_decisions_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint32,
                                              ctypes.c_uint32, _u32p, _vp, ctypes.c_uint64]
        L.tpo_render_f32.restype = ctypes.c_uint64
        L.tpo_render_f32.argtypes = [_f32p, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _f32p, _f32p, _f32p,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _f64p, ctypes.c_int]
        L.tpo_gaussian_from_ply.argtypes = [_f32p, ctypes.c_float, _f32p, _f32p, _f32p, ctypes.c_uint32, _f32p]
        L.tpo_render_f64.restype = ctypes.c_uint64
        L.tpo_render_f64.argtypes = [_f32p, ctypes.c_uint32, _vp, _vp, ctypes.c_uint32, _f32p, _f32p, _f32p,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f64p, ctypes.c_int]
        _LIB = L
    return _LIB


def ref_camera():
    """The reference's own camera/math code (oracle/_ref/libtpd_ref_camera.so) or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libtpd_ref_camera.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.tpr_look_at.argtypes = [_f32p, _f32p, _f32p, _f32p]
        R.tpr_perspective.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, _f32p]
        R.tpr_mat4_mul.argtypes = [_f32p, _f32p, _f32p]
        R.tpr_dot3.argtypes = [_f32p, _f32p, _f32p]
        R.tpr_cross3.argtypes = [_f32p, _f32p, _f32p]
        _REF = R
    return _REF


def ref_ply():
    """The reference's own PLY loader (oracle/_ref/libtpd_ref_ply.so) or None if not built."""
    path = os.path.join(_HERE, "_ref", "libtpd_ref_ply.so")
    if not os.path.exists(path):
        return None
    R = ctypes.CDLL(path)
    R.tpr_from_model.restype = ctypes.c_long
    R.tpr_from_model.argtypes = [ctypes.c_char_p, _f32p, ctypes.c_ulong]
    return R


def ref_from_model(path: str):
    """(N, 60) records from the reference's GaussianPoint::fromModel, or None when it throws or is not built."""
    R = ref_ply()
    if R is None:
        return None
    n = R.tpr_from_model(path.encode(), np.zeros(60, np.float32), 0)
    if n < 0:
        return None
    out = np.zeros((max(n, 1), 60), np.float32)
    R.tpr_from_model(path.encode(), out, n)
    return out[:n]


def gaussians_from_ply_columns(xyz, opacity, rot_wxyz, log_scale, features):
    """The oracle's restatement of fromModel's per-vertex conversion over column arrays (N x ...)."""
    L = lib()
    xyz, rot, ls, ft = _f32(xyz), _f32(rot_wxyz), _f32(log_scale), _f32(features)
    op = _f32(opacity)
    n = xyz.shape[0]
    out = np.zeros((n, 60), np.float32)
    row = np.zeros(60, np.float32)
    for i in range(n):
        L.tpo_gaussian_from_ply(xyz[i], float(op[i]), rot[i], ls[i], np.ascontiguousarray(ft[i]), ft.shape[1], row)
        out[i] = row
    return out


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nthreads_default() -> int:
    return max(1, min(os.cpu_count() or 1, 16))


# ------------------------------------------------------------------ camera

def look_at(eye, center, up) -> np.ndarray:
    v = np.zeros(16, np.float32)
    lib().tpo_look_at(_f32(eye), _f32(center), _f32(up), v)
    return v.reshape(4, 4)


def perspective(w: int, h: int, fy: float | None = None, near: float = 0.01, far: float = 100.0) -> np.ndarray:
    p = np.zeros(16, np.float32)
    if fy is None:
        fy = float(np.sqrt(np.float32(3.0)).astype(np.float32))
    lib().tpo_perspective(w, h, fy, near, far, p)
    return p.reshape(4, 4)


def fy_from_fov(deg: float) -> float:
    return float(lib().tpo_fy_from_fov(deg))


def mat4_mul(a, b) -> np.ndarray:
    o = np.zeros(16, np.float32)
    lib().tpo_mat4_mul(_f32(a).ravel(), _f32(b).ravel(), o)
    return o.reshape(4, 4)


def camera_upload(view, proj):
    pv = np.zeros(16, np.float32)
    f = np.zeros(2, np.float32)
    lib().tpo_camera_upload(_f32(view).ravel(), _f32(proj).ravel(), pv, f)
    return pv.reshape(4, 4), f


def expf(x) -> np.ndarray:
    x = _f32(x).ravel()
    y = np.empty_like(x)
    lib().tpo_expf_vec(x, y, x.size)
    return y


def expf_sweep(lo: float, hi: float, nthreads: int = 8):
    """Every float of [lo, hi] (lo < hi <= 0): (largest relative error of expf against the
    double-precision exp, number of neighbouring floats whose results decrease)."""
    m, d = ctypes.c_double(), ctypes.c_uint64()
    lib().tpo_expf_sweep(float(lo), float(hi), int(nthreads), ctypes.byref(m), ctypes.byref(d))
    return m.value, d.value


# ------------------------------------------------------------------ stages

class Camera:
    """view (4x4), proj_view (4x4), focal_ndc (2,), width, height -- the 136-B camera UBO."""

    def __init__(self, view, proj_view, focal_ndc, width, height):
        self.view = _f32(view).reshape(4, 4)
        self.proj_view = _f32(proj_view).reshape(4, 4)
        self.focal_ndc = _f32(focal_ndc).reshape(2)
        self.width = int(width)
        self.height = int(height)

    @property
    def grid(self):
        return (self.width + 15) // 16, (self.height + 15) // 16

    @property
    def ntiles(self):
        gx, gy = self.grid
        return gx * gy


def make_camera(eye, center, up, width, height, fy=None, near=0.01, far=100.0) -> Camera:
    view = look_at(eye, center, up)
    proj = perspective(width, height, fy, near, far)
    pv, f = camera_upload(view, proj)
    return Camera(view, pv, f, width, height)


def preprocess(gauss, cam: Camera, sh_degree=3, ent_idx=None, models=None, nthreads=None):
    g = _f32(gauss).reshape(-1, 60)
    n = g.shape[0]
    color = np.zeros((n, 3), np.float32)
    texel = np.zeros((n, 4), np.float32)
    copac = np.zeros((n, 4), np.float32)
    tiles = np.zeros(n, np.uint32)
    rect = np.zeros((n, 4), np.int32)
    ei = None if ent_idx is None else np.ascontiguousarray(ent_idx, np.uint32)
    md = None if models is None else _f32(models).reshape(-1, 16)
    lib().tpo_preprocess_f32(g.ravel() if n else np.zeros(60, np.float32), n, _ptr(ei), _ptr(md),
                             0 if md is None else md.shape[0], cam.view.ravel(), cam.proj_view.ravel(),
                             cam.focal_ndc, cam.width, cam.height, sh_degree, color.ravel() if n else np.zeros(3, np.float32),
                             texel.ravel() if n else np.zeros(4, np.float32), copac.ravel() if n else np.zeros(4, np.float32),
                             tiles if n else np.zeros(1, np.uint32), rect.ravel() if n else np.zeros(4, np.int32),
                             nthreads or nthreads_default())
    return dict(color=color, texel=texel, copac=copac, tiles=tiles, rect=rect)


def scan(tiles):
    t = np.ascontiguousarray(tiles, np.uint32)
    off = np.zeros(max(t.size, 1), np.uint32)
    total = lib().tpo_scan(t if t.size else np.zeros(1, np.uint32), t.size, off)
    return off[: t.size], int(total)


def duplicate(texel, offsets, total, width, height, nthreads=None):
    tx = _f32(texel).reshape(-1, 4)
    n = tx.shape[0]
    keys = np.zeros(max(total, 1), np.uint64)
    vals = np.zeros(max(total, 1), np.uint32)
    if n:
        lib().tpo_duplicate(tx.ravel(), np.ascontiguousarray(offsets, np.uint32), n, width, height, keys, vals,
                            nthreads or nthreads_default())
    return keys[:total], vals[:total]


def sort_pairs(keys, vals, nthreads=None):
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    if k.size:
        lib().tpo_sort_pairs(k, v, k.size, nthreads or nthreads_default())
    return k, v


def ranges(keys, ntiles):
    k = np.ascontiguousarray(keys, np.uint64)
    r = np.zeros((ntiles, 2), np.uint32)
    lib().tpo_ranges(k if k.size else np.zeros(1, np.uint64), k.size, ntiles, r.ravel())
    return r


def blend(splats, vals, rng, width, height, nthreads=None):
    n = splats["color"].shape[0]
    img = np.zeros((height, width, 3), np.float32)
    rgba = np.zeros((height, width, 4), np.uint8)
    v = np.ascontiguousarray(vals, np.uint32)
    lib().tpo_blend_f32(_f32(splats["color"]).ravel() if n else np.zeros(3, np.float32),
                        _f32(splats["texel"]).ravel() if n else np.zeros(4, np.float32),
                        _f32(splats["copac"]).ravel() if n else np.zeros(4, np.float32), n,
                        v if v.size else np.zeros(1, np.uint32), np.ascontiguousarray(rng, np.uint32).ravel(),
                        width, height, _ptr(img), _ptr(rgba), nthreads or nthreads_default())
    return img, rgba


def blend_decisions(splats, vals, rng, width, height):
    """Per pixel (row-major) the sorted positions k of the splats the FP32 blend used, in order.
    Returns (counts[H*W], positions[total])."""
    n = splats["color"].shape[0]
    args = [_f32(splats["color"]).ravel(), _f32(splats["texel"]).ravel(), _f32(splats["copac"]).ravel(), n,
            np.ascontiguousarray(vals, np.uint32) if len(vals) else np.zeros(1, np.uint32),
            np.ascontiguousarray(rng, np.uint32).ravel(), width, height]
    counts = np.zeros(width * height, np.uint32)
    total = lib().tpo_blend_decisions_f32(*args, counts, None, 0)
    pos = np.zeros(max(int(total), 1), np.uint32)
    lib().tpo_blend_decisions_f32(*args, counts, _ptr(pos), pos.size)
    return counts, pos[: int(total)]


def render(gauss, cam: Camera, sh_degree=3, ent_idx=None, models=None, nthreads=None, stages=False):
    """The whole FP32 frame stage by stage; returns a dict of every intermediate."""
    sp = preprocess(gauss, cam, sh_degree, ent_idx, models, nthreads)
    off, total = scan(sp["tiles"])
    keys, vals = duplicate(sp["texel"], off, total, cam.width, cam.height, nthreads)
    sk, sv = sort_pairs(keys, vals, nthreads)
    rng = ranges(sk, cam.ntiles)
    img, rgba = blend(sp, sv, rng, cam.width, cam.height, nthreads)
    return dict(splats=sp, offsets=off, total=total, keys_unsorted=keys, vals_unsorted=vals, keys=sk, vals=sv,
                ranges=rng, image=img, rgba8=rgba)


def render_timed(gauss, cam: Camera, sh_degree=3, nthreads=None):
    """The FP32 frame in one C call; returns (P, image, rgba8, per-stage seconds)."""
    g = _f32(gauss).reshape(-1, 60)
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    rgba = np.zeros((cam.height, cam.width, 4), np.uint8)
    st = np.zeros(5, np.float64)
    P = lib().tpo_render_f32(g.ravel(), g.shape[0], None, None, 0, cam.view.ravel(), cam.proj_view.ravel(),
                             cam.focal_ndc, cam.width, cam.height, sh_degree, _ptr(img), _ptr(rgba), st,
                             nthreads or nthreads_default())
    return int(P), img, rgba, st


def render_f64(gauss, cam: Camera, sh_degree=3, ent_idx=None, models=None, nthreads=None):
    """FP64 twin of the whole frame; returns (P64, image float64 HxWx3)."""
    g = _f32(gauss).reshape(-1, 60)
    img = np.zeros((cam.height, cam.width, 3), np.float64)
    ei = None if ent_idx is None else np.ascontiguousarray(ent_idx, np.uint32)
    md = None if models is None else _f32(models).reshape(-1, 16)
    P = lib().tpo_render_f64(g.ravel() if g.shape[0] else np.zeros(60, np.float32), g.shape[0], _ptr(ei), _ptr(md),
                             0 if md is None else md.shape[0], cam.view.ravel(), cam.proj_view.ravel(), cam.focal_ndc,
                             cam.width, cam.height, sh_degree, img, nthreads or nthreads_default())
    return int(P), img
