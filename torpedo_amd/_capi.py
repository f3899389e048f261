"""ctypes binding of libtpd_hip.so (declarations in include/tpd_hip.h).

There is no fallback: if the library has not been built (``python -m torpedo_amd.build``) or
cannot be loaded, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtpd_hip.so")

TPD_OK = 0
STATUS = {0: "TPD_OK", 1: "TPD_ERR_INVALID", 2: "TPD_ERR_HIP", 3: "TPD_ERR_NO_SCENE", 4: "TPD_ERR_NO_CAMERA",
          5: "TPD_ERR_CAPACITY", 6: "TPD_ERR_STATE", 7: "TPD_ERR_NO_DEVICE"}

# Every symbol include/tpd_hip.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int
_f32 = ctypes.c_float
_fp = ctypes.POINTER(ctypes.c_float)


class TpdConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("keep_unsorted", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 6)]


class TpdView(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("view", ctypes.c_float * 16), ("proj_view", ctypes.c_float * 16), ("focal_ndc", ctypes.c_float * 2)]


class TpdSceneGrads(ctypes.Structure):
    _fields_ = [("means", ctypes.c_void_p), ("log_scales", ctypes.c_void_p), ("quats", ctypes.c_void_p),
                ("opacity_logits", ctypes.c_void_p), ("sh", ctypes.c_void_p)]


SIGNATURES = {
    "tpd_last_error": (ctypes.c_char_p, []),
    "tpd_create": (_i32, [ctypes.POINTER(_vp), ctypes.POINTER(TpdConfig)]),
    "tpd_destroy": (_i32, [_vp]),
    "tpd_upload": (_i32, [_vp, _vp, _u32, _vp, _vp, _u32, _u32]),
    "tpd_bind_device_scene": (_i32, [_vp, _vp, _u32, _u32]),
    "tpd_set_transform": (_i32, [_vp, _u32, _vp]),
    "tpd_set_camera": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32]),
    "tpd_render": (_i32, [_vp, _vp]),
    "tpd_synchronize": (_i32, [_vp]),
    "tpd_get_pair_count": (_i32, [_vp, ctypes.POINTER(_u32)]),
    "tpd_read_image_f32": (_i32, [_vp, _vp]),
    "tpd_read_image_rgba8": (_i32, [_vp, _vp]),
    "tpd_read_splats": (_i32, [_vp, _vp]),
    "tpd_read_tiles": (_i32, [_vp, _vp, _vp]),
    "tpd_read_pairs": (_i32, [_vp, _vp, _vp, _u32, _i32]),
    "tpd_read_ranges": (_i32, [_vp, _vp]),
    "tpd_read_tile_order": (_i32, [_vp, _vp]),
    "tpd_device_image": (_i32, [_vp, ctypes.POINTER(_vp)]),
    "tpd_copy_image": (_i32, [_vp, _vp, _vp]),
    "tpd_present_rgba8": (_i32, [_vp, _vp, ctypes.c_uint64, _vp]),
    "tpd_save_png": (_i32, [_vp, ctypes.c_char_p]),
    "tpd_write_png": (_i32, [_vp, _u32, _u32, ctypes.c_char_p]),
    "tpd_backward": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp]),
    "tpd_read_splat_grads": (_i32, [_vp, _vp]),
    "tpd_get_backward_times": (_i32, [_vp, _vp]),
    "tpd_alloc": (_i32, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "tpd_free": (_i32, [_vp, _vp]),
    "tpd_memcpy": (_i32, [_vp, _vp, _vp, ctypes.c_uint64, _i32]),
    "tpd_set_profiling": (_i32, [_vp, _i32]),
    "tpd_get_stage_times": (_i32, [_vp, _vp, ctypes.POINTER(_u32)]),
    "tpd_get_blend_evaluations": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "tpd_measure_copy": (_i32, [_vp, ctypes.c_uint64, _u32, ctypes.POINTER(_f32)]),
    "tpd_measure_bandwidth": (_i32, [_vp, _u32, ctypes.c_uint64, _u32, _u32, ctypes.POINTER(_f32)]),
    "tpd_measure_run_scatter": (_i32, [_vp, _u32, _u32, _u32, _u32, _u32, _u32, ctypes.POINTER(_f32)]),
    "tpd_sort_pairs": (_i32, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, ctypes.POINTER(_i32)]),
    "tpd_sort_pairs_noalloc": (_i32, [_vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp, ctypes.POINTER(_i32)]),
    "tpd_autotune": (_i32, [_vp, _vp, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "tpd_set_sort_shape": (_i32, [_vp, _u32]),
    "tpd_scene_add": (_i32, [_vp, _vp, _u32, _vp, _vp, _u32, _u32, ctypes.POINTER(_u32)]),
    "tpd_scene_add_device": (_i32, [_vp, _vp, _u32, _u32, ctypes.POINTER(_u32)]),
    "tpd_scene_set_transform": (_i32, [_vp, _u32, _u32, _vp]),
    "tpd_scene_clear": (_i32, [_vp]),
    "tpd_batch_set_views": (_i32, [_vp, ctypes.POINTER(TpdView), _u32]),
    "tpd_batch_set_graph": (_i32, [_vp, _i32]),
    "tpd_batch_render": (_i32, [_vp, _i32, _vp]),
    "tpd_batch_synchronize": (_i32, [_vp]),
    "tpd_batch_backward": (_i32, [_vp, _vp, ctypes.POINTER(TpdSceneGrads), _i32, _vp]),
    "tpd_batch_step": (_i32, [_vp, _vp, ctypes.POINTER(TpdSceneGrads), _i32, _vp]),
    "tpd_batch_get_pair_count": (_i32, [_vp, ctypes.POINTER(_u32)]),
    "tpd_batch_get_item_counts": (_i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "tpd_batch_view_layout": (_i32, [_vp, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32),
                                     ctypes.POINTER(ctypes.c_uint64)]),
    "tpd_batch_read_image_f32": (_i32, [_vp, _u32, _vp]),
    "tpd_batch_read_image_rgba8": (_i32, [_vp, _u32, _vp]),
    "tpd_batch_present_rgba8": (_i32, [_vp, _u32, _vp, ctypes.c_uint64, _vp]),
    "tpd_batch_save_png": (_i32, [_vp, _u32, ctypes.c_char_p]),
    "tpd_batch_device_images": (_i32, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_uint64)]),
    "tpd_batch_read_pairs": (_i32, [_vp, _vp, _vp, _u32, ctypes.POINTER(_u32)]),
    "tpd_batch_get_times": (_i32, [_vp, _vp, ctypes.POINTER(_u32)]),
    "tpd_batch_get_sort_times": (_i32, [_vp, _vp, _vp]),
    "tpd_batch_get_blend_evaluations": (_i32, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "tpd_batch_get_sort_shapes": (_i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "tpd_batch_autotune": (_i32, [_vp, _vp, _u32, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "tpd_load_ply": (_i32, [ctypes.c_char_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)), ctypes.POINTER(_u32)]),
    "tpd_free_host": (None, [_vp]),
    "tpd_look_at": (None, [_vp, _vp, _vp, _vp]),
    "tpd_perspective": (None, [_u32, _u32, _f32, _f32, _f32, _vp]),
    "tpd_fy_from_fov": (_f32, [_f32]),
    "tpd_mat4_mul": (None, [_vp, _vp, _vp]),
    "tpd_camera_upload": (None, [_vp, _vp, _vp, _vp]),
}

_LIB = None


class TpdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


LOADED_AFTER_TORCH = False


def lib():
    """The loaded library.  If torch is used in the same process it must be imported first: torch
    carries its own HIP runtime, and the library then binds to it instead of loading a second one."""
    global _LIB, LOADED_AFTER_TORCH
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m torpedo_amd.build`")
        import sys
        LOADED_AFTER_TORCH = "torch" in sys.modules
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(code: int) -> None:
    if code != TPD_OK:
        raise TpdError(code, lib().tpd_last_error().decode())


def ptr(a: np.ndarray | None):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)
