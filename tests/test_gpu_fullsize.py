"""The benchmark workloads at their full sizes, on the GPU through the C ABI.

  multiview-32-clusters-1M   one scene of 1M clustered Gaussians seen by the 32-camera rig (radius 8,
                             resolutions 256x256 .. 1920x1080 drawn with seed 4): every view of the
                             batched frame bit-identical to its plain frame (keys, values, FP32 image,
                             RGBA8), the batch's pair total equal to the sum over the plain frames,
                             two views bit-exact against the oracle with their FP64 error stated;
  clusters-3M-1920x1080      forward bit-exact against the oracle (splats, offsets, pairs, ranges,
                             images) with the FP64 error stated; backward: finite gradients, zero
                             gradients for culled Gaussians, the plain backward equal to a one-view
                             batch bit for bit, and repeating bit for bit; FP64 finite differences and
                             the FP64 dense reference on the first
                             2,000 Gaussians of the layout moved into view, at 64x64;
  multiscene-16x4            16 scenes of 50k..2M Gaussians x 4 views of 512..1600 x 512..900, one
                             batched forward + backward step: every view's image bit-identical to its
                             plain frame, one view per scene bit-exact against the oracle with its FP64
                             error stated, and each scene's gradients equal bit for bit to the sum of
                             its plain backwards over its views.
"""
import sys

import numpy as np
import pytest

from conftest import report_fp64

pytestmark = pytest.mark.gpu

GROUPS = ("means", "log_scales", "quats", "opacity_logits", "sh")
WIDTHS = (3, 3, 4, 1, 48)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rig_view(scene_id, eye, w, h, up):
    from torpedo_amd.engine import PerspectiveCamera
    cam = PerspectiveCamera(w, h)
    cam.lookAt(list(eye), [0.0, 0.0, 0.0], list(up))
    v, pv, fo = cam.uniforms()
    return (scene_id, v, pv, fo, w, h)


@pytest.fixture(scope="module")
def engines():
    from torpedo_amd.engine import GaussianEngine
    b = GaussianEngine()
    p = GaussianEngine(keep_unsorted=True)  # the forward comparison reads the pairs before the sort too
    yield b, p
    b.close()
    p.close()


def split_view(keys, vals, tb, ib, nt):
    """One view's segment of the batch's sorted pairs, in that view's own tile and index space."""
    tiles = keys >> np.uint64(32)
    lo, hi = np.searchsorted(tiles, np.uint64(tb)), np.searchsorted(tiles, np.uint64(tb + nt))
    kv = keys[lo:hi]
    local_k = ((kv >> np.uint64(32)) - np.uint64(tb)) << np.uint64(32) | (kv & np.uint64(0xFFFFFFFF))
    return local_k, vals[lo:hi] - np.uint32(ib), hi - lo


# ------------------------------------------------------------------ 32 views of 1M Gaussians

def test_multiview_32_clusters_1m(engines, orc):
    from torpedo_amd import scenes
    beng, plain = engines
    g = scenes.cluster_scene(1_000_000, seed=2)
    views = [rig_view(0, eye, w, h, (0.0, 0.0, 1.0)) for eye, (w, h) in scenes.multiview_rig(32, 8.0, seed=4)]
    beng.clearScenes()
    beng.addScene(g)
    beng.setViews(views)
    beng.setBatchGraph(True)
    tuned = beng.autotuneBatch()  # the radix block shape the benchmark runs with
    beng.rasterBatch()
    beng.syncBatch()
    keys, vals = beng.readBatchPairs()
    P = beng.batchPairCount
    say(f"[fullsize] multiview-32: P={P}, radix shape {tuned['best_shape']}")
    assert np.all(keys[1:] >= keys[:-1])
    plain.uploadRecords(g, 3)
    total = 0
    oracle_views = {13, 6}  # a 1920x1080 view and a 256x256 one
    for vi, (sid, v, pv, fo, w, h) in enumerate(views):
        tb, ib, _ = beng.batchViewLayout(vi)
        nt = ((w + 15) // 16) * ((h + 15) // 16)
        lk, lv, cnt = split_view(keys, vals, tb, ib, nt)
        plain.setCamera(v, pv, fo, w, h)
        plain.render()
        pk, pvals = plain.readPairs()
        assert np.array_equal(lk, pk) and np.array_equal(lv, pvals), vi
        img = beng.readBatchImage(vi)
        pimg = plain.readImage()
        assert np.array_equal(img.view(np.uint32), pimg.view(np.uint32)), vi
        assert np.array_equal(beng.readBatchRGBA8(vi), plain.readRGBA8()), vi
        total += plain.pairCount
        if vi in oracle_views:
            cam = orc.Camera(v, pv, fo, w, h)
            ref = orc.render(g, cam, 3)
            assert np.array_equal(pk, ref["keys"]) and np.array_equal(pvals, ref["vals"]), vi
            assert np.array_equal(plain.readRanges(), ref["ranges"]), vi
            assert np.abs(img - ref["image"]).max() <= 1e-5, vi
            assert np.array_equal(plain.readRGBA8(), ref["rgba8"]), vi
            _, img64 = orc.render_f64(g, cam, 3)
            assert report_fp64(f"multiview-32 view {vi} ({w}x{h})", img, img64)[2] < 1e-4
        say(f"[fullsize] multiview-32 view {vi} {w}x{h}: {cnt} pairs, identical to the plain frame")
    assert total == P, (total, P)


# ------------------------------------------------------------------ 3M Gaussians at 1080p

@pytest.fixture(scope="module")
def scene_3m():
    from torpedo_amd import scenes
    return scenes.cluster_scene(3_000_000, seed=3)


def c3_camera(orc, w=1920, h=1080):
    return orc.make_camera([0.0, 0.0, -8.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0], w, h)


def test_clusters_3m_forward(engines, orc, scene_3m):
    from test_gpu_forward import compare_frame
    _, plain = engines
    compare_frame(plain, orc, scene_3m, c3_camera(orc), 3, "clusters-3M-1920x1080")


def test_clusters_3m_backward(engines, orc, scene_3m):
    beng, plain = engines
    g = scene_3m
    n = g.shape[0]
    cam = c3_camera(orc)
    W, H = cam.width, cam.height
    up = np.random.default_rng(3).standard_normal((H, W, 3)).astype(np.float32)
    d_up = plain.alloc(up.nbytes)
    plain.upload(d_up, up)
    bufs = [plain.alloc(n * k * 4) for k in WIDTHS]
    bbufs = [beng.alloc(n * k * 4) for k in WIDTHS]
    d_bup = beng.alloc(up.nbytes)
    beng.upload(d_bup, up)
    try:
        say("[fullsize] 3M backward")
        plain.uploadRecords(g, 3)
        plain.setCamera(cam.view, cam.proj_view, cam.focal_ndc, W, H)
        plain.render()
        plain.backward(d_up, *bufs)
        plain.waitIdle()
        got = []
        for b, k in zip(bufs, WIDTHS):
            h = np.empty((n, k), np.float32)
            plain.download(h, b)
            got.append(h)
        plain.backward(d_up, *bufs)  # the same backward again: bit for bit the same gradients
        plain.waitIdle()
        for name, b, k, first in zip(GROUPS, bufs, WIDTHS, got):
            again = np.empty((n, k), np.float32)
            plain.download(again, b)
            assert np.array_equal(again.view(np.uint32), first.view(np.uint32)), name
        counts, _ = plain.readTiles()
        culled = counts == 0
        say(f"[fullsize] 3M backward: {int(culled.sum())} culled of {n}, P={plain.pairCount}")
        assert 0 < culled.sum() < n
        for name, gr in zip(GROUPS, got):
            assert np.all(np.isfinite(gr)), name
            assert not gr[culled].any(), name
            assert np.abs(gr[~culled]).max() > 0, name
        # the same view as a one-view batch: the same image and the same gradients, bit for bit
        beng.clearScenes()
        beng.addScene(g)
        beng.setViews([(0, cam.view, cam.proj_view, cam.focal_ndc, W, H)])
        beng.setBatchGraph(True)
        beng.batchStep(d_bup, [bbufs])
        beng.syncBatch()
        assert np.array_equal(beng.readBatchImage(0).view(np.uint32), plain.readImage().view(np.uint32))
        for name, b, k, ref in zip(GROUPS, bbufs, WIDTHS, got):
            h = np.empty((n, k), np.float32)
            beng.download(h, b)
            d = np.linalg.norm(h.astype(np.float64) - ref) / np.linalg.norm(ref.astype(np.float64))
            say(f"[fullsize] 3M one-view batch vs plain, {name}: rel {d:.2e}")
            assert np.array_equal(h.view(np.uint32), ref.view(np.uint32)), (name, d)
    finally:
        for b in bufs + [d_up]:
            plain.free(b)
        for b in bbufs + [d_bup]:
            beng.free(b)
        beng.clearScenes()


def per_entry_report(name, g_gpu, g_ref):
    """Entry-wise error of a gradient group, |gpu - fp64| / max(|fp64|, floor * max|fp64|), for the
    floors 1e-3, 1e-2 and 1e-1; returns the maxima."""
    a = np.asarray(g_gpu, np.float64).ravel()
    b = np.asarray(g_ref, np.float64).ravel()
    out = {}
    for floor in (1e-3, 1e-2, 1e-1):
        rel = np.abs(a - b) / np.maximum(np.abs(b), floor * np.abs(b).max())
        out[floor] = float(rel.max())
    say(f"[fullsize] {name}: entry-wise rel err max (floor 1e-3 / 1e-2 / 1e-1 of the largest) "
        f"{out[1e-3]:.2e} / {out[1e-2]:.2e} / {out[1e-1]:.2e}, "
        f"norm-wise {np.linalg.norm(a - b) / np.linalg.norm(b):.2e}")
    return out


def test_clusters_3m_subset_finite_differences(engines, orc, scene_3m):
    """FP64 finite differences and the FP64 dense reference on the first 2,000 Gaussians of the
    3M layout, their means moved into the view of a camera at distance 3 (centred, scaled into
    [-1, 1]^3), at 64x64: every gradient group within 1e-4 of FP64 norm-wise; entry by entry within
    1e-3 relative (below 1% of the group's largest entry, within 1e-5 of that largest entry: FP32
    sums of terms that cancel); sampled entries within 1e-4 (of the group's largest entry) of
    central differences."""
    import torch
    from oracle import torch_ref
    from torpedo_amd.autograd import pack_records, rasterize
    from torpedo_amd.engine import PerspectiveCamera
    _, plain = engines
    g = scene_3m[:2000].copy()
    c = g[:, 0:3].mean(0)
    g[:, 0:3] = (g[:, 0:3] - c) / np.abs(g[:, 0:3] - c).max()
    W = H = 64
    p32 = [p.to(torch.float32).cuda().requires_grad_(True) for p in torch_ref.make_params(g)]
    cam = PerspectiveCamera(W, H)
    cam.lookAt([0.0, 0.0, -3.0], [0, 0, 0], [0, -1, 0])
    img = rasterize(*p32, plain, cam, 3)
    rng = np.random.default_rng(3)
    up = rng.normal(size=(H, W, 3))
    img.backward(torch.tensor(up, dtype=torch.float32, device=img.device))
    torch.cuda.synchronize()
    g_gpu = [p.grad.detach().cpu().numpy().astype(np.float64) for p in p32]
    rec = pack_records(*[p.detach() for p in p32]).cpu().numpy()
    view, pv, fo = cam.uniforms()
    ocam = orc.Camera(view, pv, fo, W, H)
    ref = orc.render(rec, ocam, 3)
    assert np.array_equal(img.detach().cpu().numpy(), ref["image"])
    _, img64 = orc.render_f64(rec, ocam, 3)
    report_fp64("clusters-3M first-2000 subset 64x64", ref["image"], img64)
    counts, pos = orc.blend_decisions(ref["splats"], ref["vals"], ref["ranges"], W, H)
    say(f"[fullsize] subset: {int(counts.sum())} blended pixel-splat pairs")
    assert counts.sum() > 2000
    params64 = [p.detach().cpu().double() for p in p32]
    g_ref, _ = torch_ref.grads(params64, ocam, 3, ref["vals"], counts, pos, up)
    entry = {name: per_entry_report(name, g_gpu[k], g_ref[k]) for k, name in enumerate(GROUPS)}
    for k, name in enumerate(GROUPS):
        assert torch_ref.relative_error(g_gpu[k], g_ref[k]) < 1e-4, name
        assert entry[name][1e-2] < 1e-3, (name, entry[name])
    used = np.unique(np.asarray(ref["vals"])[pos])
    for k, name in enumerate(GROUPS):
        width = g_gpu[k].reshape(2000, -1).shape[1]
        scale = np.abs(g_ref[k]).max()
        for gi in rng.choice(used, size=6, replace=False):
            idx = int(gi) * width + int(rng.integers(width))
            fd = torch_ref.finite_difference(params64, ocam, 3, ref["vals"], counts, pos, up, k, idx)
            assert abs(g_gpu[k].reshape(-1)[idx] - fd) <= 1e-4 * scale, (name, gi, idx, fd)


# ------------------------------------------------------------------ 16 scenes x 4 views

@pytest.fixture(scope="module")
def multiscene(engines):
    """One batched forward + backward step of the 16-scene x 4-view workload; its images and
    per-scene gradients, read back."""
    from torpedo_amd import scenes
    beng, _ = engines
    sizes = scenes.multiscene_sizes(16, 5)
    rigs = scenes.multiscene_views(16, 4, 8.0, 5)
    recs = [scenes.cluster_scene(n, seed=5 + s) for s, n in enumerate(sizes)]
    beng.clearScenes()
    for r in recs:
        beng.addScene(r)
    views = [rig_view(s, eye, w, h, (0.0, 0.0, 1.0)) for s in range(16) for eye, (w, h) in rigs[s]]
    beng.setViews(views)
    npix = [v[4] * v[5] for v in views]
    dl = np.random.default_rng(5).standard_normal(3 * sum(npix)).astype(np.float32)
    d_dl = beng.alloc(dl.nbytes)
    beng.upload(d_dl, dl)
    grads = [[beng.alloc(max(n, 1) * k * 4) for k in WIDTHS] for n in sizes]
    beng.setBatchGraph(True)
    say(f"[fullsize] multiscene: {sum(sizes)} Gaussians, {len(views)} views")
    beng.batchStep(d_dl, grads)
    beng.syncBatch()
    say(f"[fullsize] multiscene: P={beng.batchPairCount}")
    images = [beng.readBatchImage(vi) for vi in range(len(views))]
    host = []
    for s, n in enumerate(sizes):
        out = []
        for b, k in zip(grads[s], WIDTHS):
            h = np.empty((n, k), np.float32)
            beng.download(h, b)
            out.append(h)
        host.append(out)
    for gs in grads:
        for b in gs:
            beng.free(b)
    beng.free(d_dl)
    beng.clearScenes()
    offs = np.concatenate([[0], np.cumsum(npix)]) * 3
    dls = [dl[offs[i]:offs[i + 1]].reshape(views[i][5], views[i][4], 3) for i in range(len(views))]
    return dict(recs=recs, views=views, images=images, grads=host, dls=dls)


@pytest.mark.parametrize("scene", range(16))
def test_multiscene_16x4_step(engines, multiscene, scene, orc):
    """Per scene: the 4 batched images bit-identical to the plain frames, the first of them also
    bit-exact with the oracle (FP64 error stated), and the scene's gradients equal to the sum of the
    plain backwards over its views."""
    _, plain = engines
    rec = multiscene["recs"][scene]
    n = rec.shape[0]
    vis = [i for i, v in enumerate(multiscene["views"]) if v[0] == scene]
    plain.uploadRecords(rec, 3)
    bufs = [plain.alloc(max(n, 1) * k * 4) for k in WIDTHS]
    d_dl = plain.alloc(max(multiscene["dls"][i].nbytes for i in vis))
    try:
        for j, vi in enumerate(vis):
            _, v, pv, fo, w, h = multiscene["views"][vi]
            plain.setCamera(v, pv, fo, w, h)
            plain.render()
            img = plain.readImage()
            assert np.array_equal(img.view(np.uint32), multiscene["images"][vi].view(np.uint32)), (scene, vi)
            if j == 0:
                cam = orc.Camera(v, pv, fo, w, h)
                ref = orc.render(rec, cam, 3)
                k, vals = plain.readPairs()
                assert np.array_equal(k, ref["keys"]) and np.array_equal(vals, ref["vals"]), (scene, vi)
                assert np.abs(img - ref["image"]).max() <= 1e-5, (scene, vi)
                _, img64 = orc.render_f64(rec, cam, 3)
                assert report_fp64(f"multiscene-16x4 scene {scene} view {vi} ({w}x{h}, N={n})", img, img64)[2] < 1e-4
            plain.upload(d_dl, np.ascontiguousarray(multiscene["dls"][vi]))
            plain.backward(d_dl, *bufs, accumulate=j > 0)
        plain.waitIdle()
        for name, b, k, got in zip(GROUPS, bufs, WIDTHS, multiscene["grads"][scene]):
            ref = np.empty((n, k), np.float32)
            plain.download(ref, b)
            den = np.linalg.norm(ref.astype(np.float64))
            d = np.linalg.norm(got.astype(np.float64) - ref) / den if den > 0 else np.linalg.norm(got)
            ndiff = int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
            say(f"[fullsize] multiscene scene {scene} {name}: rel {d:.2e}, {ndiff} of {got.size} entries differ")
            assert ndiff == 0, (scene, name, d, ndiff)
        say(f"[fullsize] multiscene scene {scene}: N={n}, {len(vis)} views identical, gradients = sum of plain")
    finally:
        for b in bufs + [d_dl]:
            plain.free(b)
