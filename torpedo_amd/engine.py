"""Python mirror of torpedo's engine interface over the C ABI (libtpd_hip.so).

The reference's interface (C++) and its Python counterpart here:

    tpd::Scene::add(ent::group(points)) / add(point)   Scene.add(points) / Scene.add(point)   (Scene.h:19-62)
    tpd::GaussianEngine::compile(scene, settings)      GaussianEngine.compile(scene, settings) (GaussianEngine.cpp:359-399)
    tpd::GaussianEngine::rasterFrame(camera)           GaussianEngine.rasterFrame(camera)      (GaussianEngine.cpp:621-712)
    tpd::GaussianEngine::getTransformHost()            GaussianEngine.getTransformHost()       (GaussianEngine.h:24)
    tpd::TransformHost::transform(entity, mat4)        TransformHost.transform(entity, m)      (TransformHost.cpp:3-12)
    tpd::Camera::lookAt(eye, center, up)               Camera.lookAt(eye, center, up)          (Camera.cpp:3-16)
    tpd::PerspectiveCamera(w, h), setVerticalFov(deg)  PerspectiveCamera(w, h), setVerticalFov (PerspectiveCamera.{h,cpp})
    tpd::GaussianPoint / utils::rgb2sh / sh2rgb        records of 60 floats, rgb2sh / sh2rgb   (GaussianGeometry.h:10-52)

Error behaviour follows the reference: compiling an empty scene warns and keeps going
(GaussianEngine.cpp:362-365), an SH degree above 3 is clamped with a warning (:366-370),
transforming an unknown entity is ignored (TransformHost.cpp:4-6).  Failures of the device
layer raise TpdError.  The reference has no readback; this engine adds readImage/readRGBA8 and
the stage readbacks used by the parity tests.
"""
from __future__ import annotations

import ctypes
import os
import warnings
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import check, f32, lib, ptr

SH_C0 = 0.28209479177387814
GAUSSIAN_FLOATS = 60


def rgb2sh(r: float, g: float, b: float) -> np.ndarray:
    """utils::rgb2sh (GaussianGeometry.h:38-45): DC coefficients for an RGB colour, rest zero."""
    sh = np.zeros(48, np.float32)
    c0 = np.float32(SH_C0)
    sh[0] = (np.float32(r) - np.float32(0.5)) / c0
    sh[1] = (np.float32(g) - np.float32(0.5)) / c0
    sh[2] = (np.float32(b) - np.float32(0.5)) / c0
    return sh


def sh2rgb(sh) -> np.ndarray:
    """utils::sh2rgb (GaussianGeometry.h:47-52)."""
    sh = f32(sh)
    return sh[:3] * np.float32(SH_C0) + np.float32(0.5)


def gaussian_point(position, opacity, quaternion=(0.0, 0.0, 0.0, 1.0), scale=(1.0, 1.0, 1.0, 1.0), sh=None):
    """One GaussianPoint record (GaussianGeometry.h:10-30): 60 float32."""
    rec = np.zeros(GAUSSIAN_FLOATS, np.float32)
    rec[0:3] = position
    rec[3] = opacity
    rec[4:8] = quaternion
    sc = list(scale)
    rec[8:12] = sc + [1.0] * (4 - len(sc))
    rec[12:60] = np.zeros(48, np.float32) if sh is None else sh
    return rec


def from_model(ply_file) -> np.ndarray:
    """GaussianPoint::fromModel (GaussianGeometry.cpp:59-127): a 3DGS PLY file as (N, 60) records
    (tpd_load_ply: sigmoid opacity, normalised xyzw quaternion, exp scales, SH in file order)."""
    recs = ctypes.POINTER(ctypes.c_float)()
    n = ctypes.c_uint32()
    check(lib().tpd_load_ply(str(ply_file).encode(), ctypes.byref(recs), ctypes.byref(n)))
    try:
        count = int(n.value)
        out = np.ctypeslib.as_array(recs, shape=(max(count, 1) * GAUSSIAN_FLOATS,)).copy()[: count * GAUSSIAN_FLOATS]
    finally:
        lib().tpd_free_host(recs)
    return out.reshape(count, GAUSSIAN_FLOATS)


def write_png(rgba, path) -> None:
    """tpd_write_png: an (H, W, 4) uint8 RGBA image as a PNG file (8 bits per channel)."""
    img = np.ascontiguousarray(rgba, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("write_png: expected an (H, W, 4) uint8 array")
    check(lib().tpd_write_png(ptr(img), int(img.shape[1]), int(img.shape[0]), os.fsencode(path)))


class Entity(int):
    """Handle returned by Scene.add (an entt::entity in the reference)."""


class Scene:
    """Scene (Scene.h:17-54): groups of points and single points; dataAll() lists groups first, then singles
    (Scene.h:130-146), which is also the entity index order of buildEntityMap (Scene.h:154-162).

    The reference walks each kind with a single-storage EnTT view (EnTT v3.15.0, a build-time
    dependency: cmake/Dependencies.cmake:9-11).  EnTT's sparse set iterates its packed array from the
    last element to the first, and a scene only ever appends, so within each kind the most recently
    added entity comes first; groups() and singles() give that order.
    """

    def __init__(self):
        self._next = 0
        self._groups: list[tuple[Entity, np.ndarray]] = []
        self._singles: list[tuple[Entity, np.ndarray]] = []

    def add(self, points) -> Entity:
        a = f32(points)
        e = Entity(self._next)
        self._next += 1
        if a.ndim == 1:
            if a.size != GAUSSIAN_FLOATS:
                raise ValueError("a single GaussianPoint has 60 floats")
            self._singles.append((e, a.copy()))
        else:
            if a.ndim != 2 or a.shape[1] != GAUSSIAN_FLOATS:
                raise ValueError("a group of GaussianPoints is an (N, 60) array")
            self._groups.append((e, a))  # borrowed, as the reference's std::span
        return e

    def count(self) -> int:
        return len(self._singles)

    def countGroup(self) -> int:
        return len(self._groups)

    def countAll(self) -> int:
        return sum(g.shape[0] for _, g in self._groups) + len(self._singles)

    def groups(self) -> list:
        """(entity, points) of every group in view order (last added first)."""
        return self._groups[::-1]

    def singles(self) -> list:
        """(entity, point) of every single point in view order (last added first)."""
        return self._singles[::-1]

    def groupSizes(self) -> list[int]:
        return [g.shape[0] for _, g in self.groups()]

    def dataAll(self) -> np.ndarray:
        parts = [g for _, g in self.groups()] + [s[None, :] for _, s in self.singles()]
        if not parts:
            return np.zeros((0, GAUSSIAN_FLOATS), np.float32)
        return np.ascontiguousarray(np.concatenate(parts, 0), np.float32)

    def buildEntityMap(self) -> dict:
        ents = [e for e, _ in self.groups()] + [e for e, _ in self.singles()]
        return {e: i for i, e in enumerate(ents)}

    def entityIndices(self) -> np.ndarray:
        """Per-Gaussian entity index (GaussianEngine.cpp:383-388)."""
        idx = []
        k = 0
        for size in self.groupSizes():
            idx.append(np.full(size, k, np.uint32))
            k += 1
        for _ in self._singles:
            idx.append(np.full(1, k, np.uint32))
            k += 1
        return np.concatenate(idx) if idx else np.zeros(0, np.uint32)


class Camera:
    """Camera (Camera.h:11-38): a world-to-view matrix plus near/far planes."""

    def __init__(self, w2c=None):
        self._view = np.eye(4, dtype=np.float32) if w2c is None else f32(w2c).reshape(4, 4).copy()
        self._near = 0.01
        self._far = 100.0

    def lookAt(self, eye, center=None, up=None):
        """lookAt(eye, center, up) (Camera.cpp:3-16) or lookAt(R, t) (Camera.cpp:18-20)."""
        if up is None:  # (R 3x3, t)
            R = f32(eye).reshape(3, 3)
            t = f32(center).reshape(3)
            v = np.eye(4, dtype=np.float32)
            v[:3, :3] = R
            v[:3, 3] = t
            self._view = v
            return
        v = np.zeros(16, np.float32)
        e, c, u = f32(eye).reshape(3), f32(center).reshape(3), f32(up).reshape(3)
        lib().tpd_look_at(ptr(e), ptr(c), ptr(u), ptr(v))
        self._view = v.reshape(4, 4)

    def setNear(self, near: float):
        self._near = float(near)

    def setFar(self, far: float):
        self._far = float(far)

    def getViewMatrix(self) -> np.ndarray:
        return self._view.copy()

    def getProjection(self) -> np.ndarray:
        raise NotImplementedError


class PerspectiveCamera(Camera):
    """PerspectiveCamera (PerspectiveCamera.h:10-31): reversed-z projection, default fy = sqrt(3) (60 deg)."""

    def __init__(self, width: int, height: int):
        super().__init__()
        self.width = int(width)
        self.height = int(height)
        self._fy = float(np.sqrt(np.float32(3.0)))

    def setVerticalFov(self, degrees: float):
        self._fy = float(lib().tpd_fy_from_fov(float(degrees)))

    def onImageSizeChange(self, w: int, h: int):
        self.width, self.height = int(w), int(h)

    def getProjection(self) -> np.ndarray:
        p = np.zeros(16, np.float32)
        lib().tpd_perspective(self.width, self.height, self._fy, self._near, self._far, ptr(p))
        return p.reshape(4, 4)

    def uniforms(self):
        """(view, projView, focalNDC) as uploaded by updateCameraBuffer (GaussianEngine.cpp:764-775)."""
        view = f32(self._view).ravel()
        proj = f32(self.getProjection()).ravel()
        pv = np.zeros(16, np.float32)
        fo = np.zeros(2, np.float32)
        lib().tpd_camera_upload(ptr(view), ptr(proj), ptr(pv), ptr(fo))
        return view.reshape(4, 4), pv.reshape(4, 4), fo


@dataclass
class Settings:
    """GaussianEngine::Settings (GaussianEngine.h:17-21)."""
    sphericalHarmonicsDegree: int = 3

    @staticmethod
    def getDefault() -> "Settings":
        return Settings()


class TransformHost:
    """TransformHost (TransformHost.h): per-entity model matrices."""

    def __init__(self, engine: "GaussianEngine", entity_map: dict):
        self._engine = engine
        self._map = dict(entity_map)

    def transform(self, entity, matrix):
        if entity not in self._map:
            return  # TransformHost.cpp:4-6
        m = f32(matrix).reshape(16)
        check(lib().tpd_set_transform(self._engine._ctx, self._map[entity], ptr(m)))


class GaussianEngine:
    """GaussianEngine (GaussianEngine.h:15-27) over libtpd_hip.so, headless."""

    def __init__(self, device: int = 0, keep_unsorted: bool = False):
        cfg = _capi.TpdConfig()
        cfg.device = device
        cfg.keep_unsorted = 1 if keep_unsorted else 0
        h = ctypes.c_void_p()
        check(lib().tpd_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._ctx = h
        self._n = 0
        self._w = self._h = 0
        self._transform_host = None
        self._scene_ref = None

    def close(self):
        if getattr(self, "_ctx", None):
            lib().tpd_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- compile -------------------------------------------------------------------------------
    def compile(self, scene: Scene, settings: Settings | None = None):
        settings = settings or Settings.getDefault()
        n = scene.countAll()
        if n == 0:
            warnings.warn("GaussianEngine - Scene compilation warning: Could NOT find a single GaussianPoint in the scene!")
        deg = int(settings.sphericalHarmonicsDegree)
        if deg > 3:
            warnings.warn("GaussianEngine - Clamping to the maximum degrees of SH support: up to 3 degrees only")
            deg = 3
        data = scene.dataAll()
        emap = scene.buildEntityMap()
        n_ent = max(len(emap), 1)
        idx = scene.entityIndices() if len(emap) > 1 else None
        models = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (n_ent, 1))
        check(lib().tpd_upload(self._ctx, ptr(data) if n else None, n, ptr(idx), ptr(models) if idx is not None else None,
                               n_ent if idx is not None else 0, deg))
        self._n = n
        self._transform_host = TransformHost(self, emap)

    def uploadRecords(self, records, sh_degree: int = 3):
        """compile() for an (N, 60) record array with a single identity entity."""
        data = f32(records).reshape(-1, GAUSSIAN_FLOATS)
        n = data.shape[0]
        check(lib().tpd_upload(self._ctx, ptr(data) if n else None, n, None, None, 0, min(int(sh_degree), 3)))
        self._n = n
        self._transform_host = TransformHost(self, {Entity(0): 0})

    def bindDeviceScene(self, d_records: int, n: int, sh_degree: int = 3):
        """Renders from n 60-float records already in device memory (borrowed pointer)."""
        check(lib().tpd_bind_device_scene(self._ctx, ctypes.c_void_p(d_records), int(n), min(int(sh_degree), 3)))
        self._n = int(n)
        self._transform_host = TransformHost(self, {Entity(0): 0})

    def getTransformHost(self) -> TransformHost:
        return self._transform_host

    # -- frame ---------------------------------------------------------------------------------
    def setCamera(self, view, proj_view, focal_ndc, width: int, height: int):
        v, pv, fo = f32(view).reshape(16), f32(proj_view).reshape(16), f32(focal_ndc).reshape(2)
        check(lib().tpd_set_camera(self._ctx, ptr(v), ptr(pv), ptr(fo), int(width), int(height)))
        self._w, self._h = int(width), int(height)

    def rasterFrame(self, camera: PerspectiveCamera, stream=None):
        view, pv, fo = camera.uniforms()
        self.setCamera(view, pv, fo, camera.width, camera.height)
        self.render(stream)

    def render(self, stream=None):
        check(lib().tpd_render(self._ctx, None if stream is None else ctypes.c_void_p(stream)))
        self.frameCount = getattr(self, "frameCount", 0) + 1

    def copyImageTo(self, d_dst: int, stream=None):
        check(lib().tpd_copy_image(self._ctx, ctypes.c_void_p(d_dst), None if stream is None else ctypes.c_void_p(stream)))

    def presentTo(self, d_dst: int, pitch: int = 0, stream=None):
        """The frame's RGBA8 image into a device surface of row pitch `pitch` bytes (0: 4 * width),
        ordered on `stream` (recordTargetCopy, GaussianEngine.cpp:865-875)."""
        check(lib().tpd_present_rgba8(self._ctx, ctypes.c_void_p(d_dst), int(pitch),
                                      None if stream is None else ctypes.c_void_p(stream)))

    def savePNG(self, path: str):
        """Writes the frame's RGBA8 image to a PNG file."""
        check(lib().tpd_save_png(self._ctx, os.fsencode(path)))

    def backward(self, d_dL_dimage: int, d_means: int, d_log_scales: int, d_quats: int, d_opacity_logits: int,
                 d_sh: int, accumulate: bool = False, stream=None):
        """tpd_backward on device pointers (gradients of the last rendered frame)."""
        vp = ctypes.c_void_p
        check(lib().tpd_backward(self._ctx, vp(d_dL_dimage), vp(d_means), vp(d_log_scales), vp(d_quats),
                                 vp(d_opacity_logits), vp(d_sh), 1 if accumulate else 0,
                                 None if stream is None else vp(stream)))

    def backwardTimes(self) -> dict:
        ms = np.zeros(3, np.float32)
        check(lib().tpd_get_backward_times(self._ctx, ptr(ms)))
        return {"splat_grad_pass": float(ms[0]), "param_grad_pass": float(ms[1]), "fwd_bwd_span": float(ms[2])}

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        check(lib().tpd_alloc(self._ctx, int(nbytes), ctypes.byref(p)))
        return int(p.value)

    def free(self, d: int):
        check(lib().tpd_free(self._ctx, ctypes.c_void_p(d)))

    def upload(self, d: int, host: np.ndarray):
        h = np.ascontiguousarray(host)
        check(lib().tpd_memcpy(self._ctx, ctypes.c_void_p(d), ptr(h), h.nbytes, 0))

    def download(self, host: np.ndarray, d: int):
        check(lib().tpd_memcpy(self._ctx, ptr(host), ctypes.c_void_p(d), host.nbytes, 1))

    def readSplatGrads(self) -> np.ndarray:
        out = np.empty((max(self._n, 1), 9), np.float32)
        check(lib().tpd_read_splat_grads(self._ctx, ptr(out)))
        return out[: self._n]

    def waitIdle(self):
        """Engine::waitIdle (Engine.h:19)."""
        check(lib().tpd_synchronize(self._ctx))

    # -- readbacks -----------------------------------------------------------------------------
    @property
    def pairCount(self) -> int:
        v = ctypes.c_uint32()
        check(lib().tpd_get_pair_count(self._ctx, ctypes.byref(v)))
        return int(v.value)

    def readImage(self) -> np.ndarray:
        img = np.empty((self._h, self._w, 3), np.float32)
        check(lib().tpd_read_image_f32(self._ctx, ptr(img)))
        return img

    def readRGBA8(self) -> np.ndarray:
        img = np.empty((self._h, self._w, 4), np.uint8)
        check(lib().tpd_read_image_rgba8(self._ctx, ptr(img)))
        return img

    def readSplats(self) -> dict:
        raw = np.empty((max(self._n, 1), 12), np.float32)
        check(lib().tpd_read_splats(self._ctx, ptr(raw)))
        raw = raw[: self._n]
        return dict(color=raw[:, 0:3].copy(), tiles=raw[:, 3].copy().view(np.uint32), texel=raw[:, 4:8].copy(),
                    copac=raw[:, 8:12].copy())

    def readTiles(self):
        counts = np.empty(max(self._n, 1), np.uint32)
        offs = np.empty(max(self._n, 1), np.uint32)
        check(lib().tpd_read_tiles(self._ctx, ptr(counts), ptr(offs)))
        return counts[: self._n], offs[: self._n]

    def readPairs(self, unsorted: bool = False):
        P = self.pairCount
        k = np.empty(max(P, 1), np.uint64)
        v = np.empty(max(P, 1), np.uint32)
        check(lib().tpd_read_pairs(self._ctx, ptr(k), ptr(v), max(P, 1), 1 if unsorted else 0))
        return k[:P], v[:P]

    def readRanges(self) -> np.ndarray:
        gx, gy = (self._w + 15) // 16, (self._h + 15) // 16
        r = np.empty((gx * gy, 2), np.uint32)
        check(lib().tpd_read_ranges(self._ctx, ptr(r)))
        return r

    def readTileOrder(self) -> np.ndarray:
        """The blend's launch order of the tiles (heaviest class of pair count first)."""
        gx, gy = (self._w + 15) // 16, (self._h + 15) // 16
        o = np.empty(gx * gy, np.uint32)
        check(lib().tpd_read_tile_order(self._ctx, ptr(o)))
        return o

    def autotune(self) -> dict:
        """Times every radix-pass block shape on the last frame's pairs and keeps the fastest."""
        ms = np.zeros(64, np.float32)
        ns = ctypes.c_uint32()
        best = ctypes.c_uint32()
        check(lib().tpd_autotune(self._ctx, ptr(ms), 64, ctypes.byref(ns), ctypes.byref(best)))
        return {"sort_ms_per_shape": [round(float(x), 4) for x in ms[: ns.value]], "best_shape": int(best.value)}

    def setSortShape(self, shape: int):
        check(lib().tpd_set_sort_shape(self._ctx, int(shape)))

    def setProfiling(self, enable: bool = True):
        check(lib().tpd_set_profiling(self._ctx, 1 if enable else 0))

    STAGES = ("preprocess_scan_duplicate", "sort_histogram", "sort_passes", "ranges", "blend", "frame")

    def stageTimes(self) -> dict:
        """Milliseconds per stage of the last profiled frame (HIP events on the render stream)."""
        ms = np.zeros(6, np.float32)
        passes = ctypes.c_uint32()
        check(lib().tpd_get_stage_times(self._ctx, ptr(ms), ctypes.byref(passes)))
        out = {k: float(v) for k, v in zip(self.STAGES, ms)}
        out["radix_passes"] = int(passes.value)
        return out

    def blendEvaluations(self) -> int:
        """Pixel-splat pairs whose power the last rendered frame's blend evaluates
        (tpd_get_blend_evaluations: the frame is blended once more by a counting kernel)."""
        ev = ctypes.c_uint64()
        check(lib().tpd_get_blend_evaluations(self._ctx, ctypes.byref(ev)))
        return int(ev.value)

    def measureCopy(self, nbytes: int = 1 << 30, iters: int = 10) -> float:
        """GB/s of a plain device copy (read + write bytes counted)."""
        ms = ctypes.c_float()
        check(lib().tpd_measure_copy(self._ctx, int(nbytes), int(iters), ctypes.byref(ms)))
        return 2.0 * nbytes / (ms.value * 1e-3) / 1e9

    BANDWIDTH_KINDS = ("copy", "memcpy", "copy_nontemporal", "read", "write")
    GATHER_KINDS = {"gather8": 5, "gather4": 6}

    def measureBandwidth(self, kind: str, nbytes: int, iters: int = 20, blocks: int = 0) -> float:
        """GB/s of one streaming pattern over nbytes (tpd_measure_bandwidth); copies count the bytes
        read and the bytes written."""
        k = self.BANDWIDTH_KINDS.index(kind)
        ms = ctypes.c_float()
        check(lib().tpd_measure_bandwidth(self._ctx, k, int(nbytes), int(iters), int(blocks), ctypes.byref(ms)))
        moved = 2.0 * nbytes if k <= 2 else float(nbytes)
        return moved / (ms.value * 1e-3) / 1e9

    def measureGather(self, kind: str, nbytes: int, iters: int = 10) -> float:
        """Random reads per second (G/s): 2^25 random 8-byte ("gather8") or 4-byte ("gather4") reads
        from an array of nbytes (tpd_measure_bandwidth kinds 5 and 6)."""
        ms = ctypes.c_float()
        check(lib().tpd_measure_bandwidth(self._ctx, self.GATHER_KINDS[kind], int(nbytes), int(iters), 0,
                                          ctypes.byref(ms)))
        return (1 << 25) / (ms.value * 1e-3) / 1e9

    def measureRunScatter(self, key_bytes: int, n: int, bins: int, keys_per_block: int, blocks_per_cu: int = 0,
                          iters: int = 20) -> float:
        """GB/s (bytes read + written) of a radix pass's memory pattern without its ranking
        (tpd_measure_run_scatter): n keys of key_bytes and n 4-byte values, blocks of keys_per_block
        keys with random digits among `bins`, written into per-(digit, block) runs; blocks_per_cu
        caps the resident blocks per CU (0: no cap)."""
        ms = ctypes.c_float()
        check(lib().tpd_measure_run_scatter(self._ctx, int(key_bytes), int(n), int(bins), int(keys_per_block),
                                            int(blocks_per_cu), int(iters), ctypes.byref(ms)))
        return 2.0 * n * (key_bytes + 4) / (ms.value * 1e-3) / 1e9

    def deviceImagePtr(self) -> int:
        p = ctypes.c_void_p()
        check(lib().tpd_device_image(self._ctx, ctypes.byref(p)))
        return int(p.value or 0)

    # -- batched rendering ---------------------------------------------------------------------
    # The reference renders one camera of one compiled scene per rasterFrame (GaussianEngine.h:26);
    # a batch holds several scenes and renders many views of them per launch (tpd_batch_*).
    def addScene(self, scene, settings: Settings | None = None) -> int:
        """Adds a Scene (or an (N, 60) record array) to the batch; returns its scene id."""
        sid = ctypes.c_uint32()
        if isinstance(scene, Scene):
            deg = min(int((settings or Settings.getDefault()).sphericalHarmonicsDegree), 3)
            data = scene.dataAll()
            emap = scene.buildEntityMap()
            n_ent = max(len(emap), 1)
            idx = scene.entityIndices() if len(emap) > 1 else None
            models = np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (n_ent, 1)) if idx is not None else None
            n = scene.countAll()
            check(lib().tpd_scene_add(self._ctx, ptr(data) if n else None, n, ptr(idx), ptr(models),
                                      n_ent if idx is not None else 0, deg, ctypes.byref(sid)))
            self._batch_maps = getattr(self, "_batch_maps", {})
            self._batch_maps[int(sid.value)] = dict(emap)
        else:
            deg = min(int(settings.sphericalHarmonicsDegree), 3) if settings else 3
            data = f32(scene).reshape(-1, GAUSSIAN_FLOATS)
            n = data.shape[0]
            check(lib().tpd_scene_add(self._ctx, ptr(data) if n else None, n, None, None, 0, deg, ctypes.byref(sid)))
        self._batch_sizes = getattr(self, "_batch_sizes", []) + [n]
        return int(sid.value)

    def addDeviceScene(self, d_records: int, n: int, sh_degree: int = 3) -> int:
        """Adds n records already in device memory (borrowed, 16-byte aligned) to the batch."""
        sid = ctypes.c_uint32()
        check(lib().tpd_scene_add_device(self._ctx, ctypes.c_void_p(d_records), int(n), min(int(sh_degree), 3),
                                         ctypes.byref(sid)))
        self._batch_sizes = getattr(self, "_batch_sizes", []) + [int(n)]
        return int(sid.value)

    def setSceneTransform(self, scene_id: int, entity, matrix):
        """TransformHost::transform for an entity of batch scene scene_id."""
        emap = getattr(self, "_batch_maps", {}).get(int(scene_id), {})
        if entity not in emap:
            return  # TransformHost.cpp:4-6
        m = f32(matrix).reshape(16)
        check(lib().tpd_scene_set_transform(self._ctx, int(scene_id), emap[entity], ptr(m)))

    def clearScenes(self):
        check(lib().tpd_scene_clear(self._ctx))
        self._batch_sizes, self._batch_views, self._batch_maps = [], [], {}

    def setViews(self, views):
        """views: list of (scene_id, PerspectiveCamera) or (scene_id, view, proj_view, focal_ndc, w, h)."""
        arr = (_capi.TpdView * max(len(views), 1))()
        sizes = []
        for i, v in enumerate(views):
            if len(v) == 2:
                sid, cam = v
                view, pv, fo = cam.uniforms()
                w, h = cam.width, cam.height
            else:
                sid, view, pv, fo, w, h = v
            arr[i].scene = int(sid)
            arr[i].width, arr[i].height = int(w), int(h)
            arr[i].view[:] = [float(x) for x in f32(view).reshape(16)]
            arr[i].proj_view[:] = [float(x) for x in f32(pv).reshape(16)]
            arr[i].focal_ndc[:] = [float(x) for x in f32(fo).reshape(2)]
            sizes.append((int(w), int(h)))
        check(lib().tpd_batch_set_views(self._ctx, arr, len(views)))
        self._batch_views = sizes

    def setBatchGraph(self, enabled: bool = True):
        check(lib().tpd_batch_set_graph(self._ctx, 1 if enabled else 0))

    def rasterBatch(self, for_backward: bool = False, stream=None):
        """Launches one batch frame (asynchronous; syncBatch or any batch read waits for it)."""
        check(lib().tpd_batch_render(self._ctx, 1 if for_backward else 0,
                                     None if stream is None else ctypes.c_void_p(stream)))

    def syncBatch(self):
        check(lib().tpd_batch_synchronize(self._ctx))

    @staticmethod
    def _grads_table(grads):
        t = (_capi.TpdSceneGrads * max(len(grads), 1))()
        for i, g in enumerate(grads):
            t[i].means, t[i].log_scales, t[i].quats, t[i].opacity_logits, t[i].sh = [int(x) for x in g]
        return t

    def batchBackward(self, d_dL_dimages: int, grads, accumulate: bool = False, stream=None):
        """grads: per scene (d_means, d_log_scales, d_quats, d_opacity_logits, d_sh) device pointers."""
        check(lib().tpd_batch_backward(self._ctx, ctypes.c_void_p(d_dL_dimages), self._grads_table(grads),
                                       1 if accumulate else 0, None if stream is None else ctypes.c_void_p(stream)))

    def batchStep(self, d_dL_dimages: int, grads, accumulate: bool = False, stream=None):
        """Forward + backward of the batch in one recorded graph (asynchronous)."""
        check(lib().tpd_batch_step(self._ctx, ctypes.c_void_p(d_dL_dimages), self._grads_table(grads),
                                   1 if accumulate else 0, None if stream is None else ctypes.c_void_p(stream)))

    @property
    def batchPairCount(self) -> int:
        v = ctypes.c_uint32()
        check(lib().tpd_batch_get_pair_count(self._ctx, ctypes.byref(v)))
        return int(v.value)

    def batchItemCounts(self) -> tuple:
        """(items, kept) of the last batch frame: all (Gaussian, view) items, and the visible ones
        that its item sort kept (tpd_batch_get_item_counts)."""
        items, kept = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().tpd_batch_get_item_counts(self._ctx, ctypes.byref(items), ctypes.byref(kept)))
        return int(items.value), int(kept.value)

    def batchViewLayout(self, view: int):
        tb, ib, pb = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().tpd_batch_view_layout(self._ctx, int(view), ctypes.byref(tb), ctypes.byref(ib), ctypes.byref(pb)))
        return int(tb.value), int(ib.value), int(pb.value)

    def readBatchImage(self, view: int) -> np.ndarray:
        w, h = self._batch_views[view]
        img = np.empty((h, w, 3), np.float32)
        check(lib().tpd_batch_read_image_f32(self._ctx, int(view), ptr(img)))
        return img

    def readBatchRGBA8(self, view: int) -> np.ndarray:
        w, h = self._batch_views[view]
        img = np.empty((h, w, 4), np.uint8)
        check(lib().tpd_batch_read_image_rgba8(self._ctx, int(view), ptr(img)))
        return img

    def presentBatchViewTo(self, view: int, d_dst: int, pitch: int = 0, stream=None):
        check(lib().tpd_batch_present_rgba8(self._ctx, int(view), ctypes.c_void_p(d_dst), int(pitch),
                                            None if stream is None else ctypes.c_void_p(stream)))

    def saveBatchPNG(self, view: int, path: str):
        check(lib().tpd_batch_save_png(self._ctx, int(view), os.fsencode(path)))

    def batchImagesPtr(self):
        p, n = ctypes.c_void_p(), ctypes.c_uint64()
        check(lib().tpd_batch_device_images(self._ctx, ctypes.byref(p), ctypes.byref(n)))
        return int(p.value or 0), int(n.value)

    def readBatchPairs(self):
        P = self.batchPairCount
        k = np.empty(max(P, 1), np.uint64)
        v = np.empty(max(P, 1), np.uint32)
        n = ctypes.c_uint32()
        check(lib().tpd_batch_read_pairs(self._ctx, ptr(k), ptr(v), max(P, 1), ctypes.byref(n)))
        return k[:P], v[:P]

    BATCH_STAGES = ("frame", "items", "item_sort", "emit", "pair_sort", "ranges", "blend")

    def batchTimes(self) -> dict:
        """GPU milliseconds of the last profiled batch frame (stages only for eager forward frames)."""
        ms = np.zeros(7, np.float32)
        passes = ctypes.c_uint32()
        check(lib().tpd_batch_get_times(self._ctx, ptr(ms), ctypes.byref(passes)))
        out = {k: float(v) for k, v in zip(self.BATCH_STAGES, ms) if v >= 0}
        out["radix_passes"] = int(passes.value)
        return out

    def batchSortTimes(self) -> dict:
        """The last eager profiled forward batch frame's two sorts (tpd_batch_get_sort_times)."""
        ms = np.zeros(4, np.float32)
        n = np.zeros(4, np.uint32)
        check(lib().tpd_batch_get_sort_times(self._ctx, ptr(ms), ptr(n)))
        return {"item_hist_ms": float(ms[0]), "item_passes_ms": float(ms[1]), "pair_hist_ms": float(ms[2]),
                "pair_passes_ms": float(ms[3]), "items": int(n[0]), "item_passes": int(n[1]), "pairs": int(n[2]),
                "pair_passes": int(n[3])}

    def batchBlendEvaluations(self) -> int:
        """Pixel-splat pairs whose power the last batch frame's blend evaluated
        (tpd_batch_get_blend_evaluations: the frame is blended once more by a counting kernel)."""
        ev = ctypes.c_uint64()
        check(lib().tpd_batch_get_blend_evaluations(self._ctx, ctypes.byref(ev)))
        return int(ev.value)

    def autotuneBatch(self) -> dict:
        ms = np.zeros(64, np.float32)
        ns, best = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().tpd_batch_autotune(self._ctx, ptr(ms), 64, ctypes.byref(ns), ctypes.byref(best)))
        item, pair = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().tpd_batch_get_sort_shapes(self._ctx, ctypes.byref(item), ctypes.byref(pair)))
        return {"sort_ms_per_shape": [round(float(x), 4) for x in ms[: ns.value]], "best_shape": int(best.value),
                "item_shape": int(item.value), "pair_shape": int(pair.value)}


def sort_pairs_device(engine: GaussianEngine, keys_ptr: int, vals_ptr: int, keys_alt_ptr: int, vals_alt_ptr: int,
                      n: int, key_bits: int, stream: int | None = None) -> bool:
    """tpd_sort_pairs on device buffers (e.g. torch tensors' data_ptr()); returns True if the result is in *_alt."""
    in_alt = ctypes.c_int()
    check(lib().tpd_sort_pairs(engine._ctx, ctypes.c_void_p(keys_ptr), ctypes.c_void_p(vals_ptr),
                               ctypes.c_void_p(keys_alt_ptr), ctypes.c_void_p(vals_alt_ptr), int(n), int(key_bits),
                               None if stream is None else ctypes.c_void_p(stream), ctypes.byref(in_alt)))
    return bool(in_alt.value)


def sort_pairs_device_noalloc(engine: GaussianEngine, keys_ptr: int, vals_ptr: int, keys_alt_ptr: int,
                              vals_alt_ptr: int, n: int, key_bits: int, stream: int | None = None) -> bool:
    """tpd_sort_pairs_noalloc: as sort_pairs_device, with no device allocation (n <= 4,194,304; the
    library's fixed per-GPU scratch, shared by every call on that GPU)."""
    in_alt = ctypes.c_int()
    check(lib().tpd_sort_pairs_noalloc(engine._ctx, ctypes.c_void_p(keys_ptr), ctypes.c_void_p(vals_ptr),
                                       ctypes.c_void_p(keys_alt_ptr), ctypes.c_void_p(vals_alt_ptr), int(n),
                                       int(key_bits), None if stream is None else ctypes.c_void_p(stream),
                                       ctypes.byref(in_alt)))
    return bool(in_alt.value)
