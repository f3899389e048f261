"""Presentation on the GPU (through the C ABI): the RGBA8 target copied into a pitched device
surface (tpd_present_rgba8, the reference's recordTargetCopy) and written as a PNG file
(tpd_save_png), for a plain frame and for views of a batch.  Both must hold exactly the bytes of
the RGBA8 readback, which test_gpu_forward.py pins to the oracle; the pitch padding must be left
untouched."""
import sys

import numpy as np
import pytest

from conftest import c1_camera, small_scene
from test_present import read_png

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def surface(h, pitch):
    import torch
    return torch.full((h, pitch), 0xA5, dtype=torch.uint8, device="cuda")


def check_surface(s, rgba, w):
    got = s.cpu().numpy()
    assert np.array_equal(got[:, :4 * w].reshape(rgba.shape), rgba)
    assert np.all(got[:, 4 * w:] == 0xA5), "pitch padding overwritten"


def test_present_and_png_plain_frame(orc, tmp_path):
    import torch
    from torpedo_amd.engine import GaussianEngine
    e = GaussianEngine()
    try:
        g = small_scene(20000, seed=41, log_scale_mean=-3.5)
        cam = c1_camera(orc, 333, 197)
        say(f"[present] plain frame {cam.width}x{cam.height}")
        e.uploadRecords(g, 3)
        e.setCamera(cam.view, cam.proj_view, cam.focal_ndc, cam.width, cam.height)
        e.render()
        rgba = e.readRGBA8()
        assert rgba.any()
        for pitch in (0, 4 * cam.width + 52):
            s = surface(cam.height, pitch or 4 * cam.width)
            torch.cuda.synchronize()
            e.presentTo(s.data_ptr(), pitch)
            e.waitIdle()
            check_surface(s, rgba, cam.width)
        path = tmp_path / "plain.png"
        e.savePNG(path)
        assert np.array_equal(read_png(path), rgba)
    finally:
        e.close()


def test_present_and_png_batch_views(tmp_path):
    import torch
    from torpedo_amd.engine import GaussianEngine, PerspectiveCamera
    e = GaussianEngine()
    try:
        g = small_scene(20000, seed=42, log_scale_mean=-3.5)
        sid = e.addScene(g)
        views = []
        for i, (w, h) in enumerate([(256, 160), (97, 301), (640, 360)]):
            cam = PerspectiveCamera(w, h)
            cam.lookAt([0.8 * i - 0.8, 0.2, -3.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0])
            v, pv, fo = cam.uniforms()
            views.append((sid, v, pv, fo, w, h))
        e.setViews(views)
        e.rasterBatch()
        e.syncBatch()
        for vi, (_, _, _, _, w, h) in enumerate(views):
            say(f"[present] batch view {vi}: {w}x{h}")
            rgba = e.readBatchRGBA8(vi)
            s = surface(h, 4 * w + 12)
            torch.cuda.synchronize()
            e.presentBatchViewTo(vi, s.data_ptr(), 4 * w + 12)
            torch.cuda.synchronize()
            e.waitIdle()
            check_surface(s, rgba, w)
            path = tmp_path / f"view{vi}.png"
            e.saveBatchPNG(vi, path)
            assert np.array_equal(read_png(path), rgba)
    finally:
        e.close()
