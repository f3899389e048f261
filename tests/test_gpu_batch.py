"""Batched rendering on the GPU (through the C ABI): every view of a batch must equal the plain
frame of the same scene and camera -- per-view sorted keys (global tile - the view's first tile)
and values (item - the view's first item), FP32 image and RGBA8 image all bit-exact -- and the
plain frame is itself pinned bit-exact to the CPU oracle by test_gpu_forward.py; one view per
case is also compared with the oracle directly.  Batched gradients equal the sum over the scene's
views of the plain backward bit for bit: the per-(tile, splat) sums are order-independent (float64
LDS accumulation of FP32 row sums) and the batch adds the views' terms in the order the plain
backward accumulates them.
"""
import sys

import numpy as np
import pytest

from conftest import small_scene

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a, np.float32).view(np.uint32),
                          np.ascontiguousarray(b, np.float32).view(np.uint32))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def engines():
    from torpedo_amd.engine import GaussianEngine
    b = GaussianEngine()
    p = GaussianEngine()
    yield b, p
    b.close()
    p.close()


def view_tuple(scene_id, eye, w, h, up=(0.0, -1.0, 0.0)):
    from torpedo_amd.engine import PerspectiveCamera
    cam = PerspectiveCamera(w, h)
    cam.lookAt(list(eye), [0.0, 0.0, 0.0], list(up))
    v, pv, fo = cam.uniforms()
    return (scene_id, v, pv, fo, w, h)


def plain_frame(plain, view):
    _, v, pv, fo, w, h = view
    plain.setCamera(v, pv, fo, w, h)
    plain.render()
    k, vals = plain.readPairs()
    return k, vals, plain.readImage(), plain.readRGBA8()


def check_batch(beng, plain, records, degs, views, name, orc=None):
    """records[s], degs[s]: scene s as added to beng; compares every view with the plain frame."""
    keys, vals = beng.readBatchPairs()
    P = beng.batchPairCount
    say(f"[batch] {name}: {len(views)} views, P={P}")
    assert keys.shape[0] == P
    tiles = (keys >> np.uint64(32)).astype(np.int64)
    assert np.all(np.diff(tiles) >= 0), "global tiles not sorted"
    covered = visible = 0
    for vi, view in enumerate(views):
        sid, _, _, _, w, h = view
        tb, ib, pb = beng.batchViewLayout(vi)
        nt = ((w + 15) // 16) * ((h + 15) // 16)
        lo, hi = np.searchsorted(tiles, tb), np.searchsorted(tiles, tb + nt)
        kv, vv = keys[lo:hi], vals[lo:hi]
        local_k = ((kv >> np.uint64(32)) - np.uint64(tb)) << np.uint64(32) | (kv & np.uint64(0xFFFFFFFF))
        local_v = vv - np.uint32(ib)
        plain.uploadRecords(records[sid], degs[sid])
        pk, pv, pimg, prgba = plain_frame(plain, view)
        visible += int(np.count_nonzero(plain.readTiles()[0]))
        assert np.array_equal(local_k, pk), (name, vi, "keys")
        assert np.array_equal(local_v, pv), (name, vi, "values")
        img = beng.readBatchImage(vi)
        assert np.array_equal(img.view(np.uint32), pimg.view(np.uint32)), (name, vi, "image")
        assert np.array_equal(beng.readBatchRGBA8(vi), prgba), (name, vi, "rgba8")
        covered += hi - lo
        if orc is not None and vi == 0:
            _, v, pvm, fo, w, h = view
            ref = orc.render(records[sid], orc.Camera(v, pvm, fo, w, h), degs[sid])
            assert np.array_equal(local_k, ref["keys"]) and np.array_equal(local_v, ref["vals"]), (name, "oracle")
            assert np.abs(img - ref["image"]).max() <= 1e-5, (name, "oracle image")
    assert covered == P, (name, covered, P)
    # the item sort keeps exactly the visible items: those with tiles in the plain frames
    items, kept = beng.batchItemCounts()
    assert items == sum(records[v[0]].shape[0] for v in views), (name, "items")
    assert kept == visible, (name, "kept items", kept, visible)


def test_batch_views_of_one_scene(engines, orc):
    from torpedo_amd import scenes
    beng, plain = engines
    beng.clearScenes()
    g = scenes.cluster_scene(30_000, seed=7)
    sid = beng.addScene(g)
    sizes = [(320, 240), (257, 131), (64, 64), (512, 288), (100, 700), (1, 1), (1280, 720)]
    eyes = scenes.fibonacci_sphere(len(sizes), 8.0)
    views = [view_tuple(sid, eyes[i], w, h, up=(0.0, 0.0, 1.0)) for i, (w, h) in enumerate(sizes)]
    beng.setViews(views)
    for graph in (True, False, True):
        beng.setBatchGraph(graph)
        beng.rasterBatch()
        beng.syncBatch()
        check_batch(beng, plain, [g], [3], views, f"one-scene graph={graph}", orc=orc if graph else None)
    from test_gpu_sort import NUM_SHAPES
    try:
        for shape in range(NUM_SHAPES):  # both sorts of the batch frame under every radix block shape
            beng.setSortShape(shape)
            beng.rasterBatch()
            beng.syncBatch()
            check_batch(beng, plain, [g], [3], views, f"one-scene shape {shape}")
    finally:
        beng.setSortShape(0)


def test_batch_emission_of_large_footprints(engines, orc):
    """Pair emission stages the pairs of items with at most four tiles in LDS, chunk by chunk of a
    block's output range (4,096 sorted items per block, chunks of 7,680 pairs), and writes them
    directly when the block's larger items push that range past four chunks (30,720 pairs).  Small
    and large splats mixed in depth order exercise both paths and chunk boundaries that cut through
    items; every view must equal its plain frame."""
    beng, plain = engines
    beng.clearScenes()
    rng = np.random.default_rng(5)
    parts = [small_scene(5000, seed=81), small_scene(1500, seed=82, log_scale_mean=-2.8),
             small_scene(600, seed=83, log_scale_mean=-1.6)]
    g = np.concatenate(parts)[rng.permutation(7100)].astype(np.float32)
    sid = beng.addScene(g)
    sizes = [(640, 480), (1280, 720), (200, 150)]
    eyes = [(0.0, 0.0, -3.2), (2.0, 0.5, -2.5), (-2.5, 1.0, 1.5)]
    views = [view_tuple(sid, eyes[i], w, h) for i, (w, h) in enumerate(sizes)]
    beng.setViews(views)
    beng.setBatchGraph(True)
    beng.rasterBatch()
    beng.syncBatch()
    P = beng.batchPairCount
    say(f"[batch] large footprints: {P} pairs over {g.shape[0] * len(views)} items")
    blocks = -(-g.shape[0] * len(views) // 4096)  # emission blocks of 4,096 depth-sorted items
    assert P > 30720 * blocks  # so that some block's range is longer than four chunks
    check_batch(beng, plain, [g], [3], views, "large footprints", orc=orc)


def test_batch_blend_evaluations_equal_the_plain_frames(engines):
    """The batch blend's evaluation count (the FLOP count behind its roofline figure) equals the sum
    of the plain frames' counts over the views, and counting leaves the batch images as they were."""
    from torpedo_amd import scenes
    beng, plain = engines
    beng.clearScenes()
    g = scenes.cluster_scene(30_000, seed=11)
    sid = beng.addScene(g)
    sizes = [(320, 240), (257, 131), (640, 360), (96, 512)]
    eyes = scenes.fibonacci_sphere(len(sizes), 8.0)
    views = [view_tuple(sid, eyes[i], w, h, up=(0.0, 0.0, 1.0)) for i, (w, h) in enumerate(sizes)]
    beng.setViews(views)
    beng.setBatchGraph(True)
    beng.rasterBatch()
    beng.syncBatch()
    before = [beng.readBatchImage(i) for i in range(len(views))]
    ev = beng.batchBlendEvaluations()
    for i in range(len(views)):
        assert same_bits(beng.readBatchImage(i), before[i]), i
    plain.uploadRecords(g, 3)
    plain.setProfiling(True)
    try:
        total = 0
        for view in views:
            plain_frame(plain, view)
            total += plain.blendEvaluations()
    finally:
        plain.setProfiling(False)
    say(f"[batch] blend evaluations: batch {ev}, plain frames {total}")
    assert ev == total and ev > 0


def test_batch_uneven_scenes(engines):
    beng, plain = engines
    beng.clearScenes()
    sizes = [0, 1, 255, 257, 5000, 20000]
    degs = [3, 0, 1, 2, 3, 3]
    recs = [small_scene(n, seed=40 + i, spread=1.0, log_scale_mean=-3.5) for i, n in enumerate(sizes)]
    from torpedo_amd.engine import Settings
    ids = []
    for r, d in zip(recs, degs):
        s = Settings()
        s.sphericalHarmonicsDegree = d
        ids.append(beng.addScene(r, s))
    # scene 3 has no view; views interleave scenes
    plan = [(5, (200, 120)), (0, (64, 48)), (4, (333, 77)), (1, (16, 16)), (5, (96, 96)), (2, (128, 256)),
            (4, (640, 360)), (5, (31, 33))]
    views = [view_tuple(ids[s], (0.3 * k, -0.2, -3.0 + 0.1 * k), w, h) for k, (s, (w, h)) in enumerate(plan)]
    beng.setViews(views)
    beng.setBatchGraph(True)
    beng.rasterBatch()
    check_batch(beng, plain, recs, degs, views, "uneven scenes")


def test_batch_camera_update_and_growth(engines):
    beng, plain = engines
    beng.clearScenes()
    g = small_scene(20_000, seed=9, log_scale_mean=-3.5)  # means in [-1, 1]^3
    sid = beng.addScene(g)
    far = [view_tuple(sid, (0.0, 0.0, -20.0 - 5 * i), w, h) for i, (w, h) in enumerate([(256, 256), (400, 300)])]
    near = [view_tuple(sid, (0.2 * i, 0.0, -2.5), w, h) for i, (w, h) in enumerate([(256, 256), (400, 300)])]
    beng.setBatchGraph(True)
    beng.setViews(far)
    beng.rasterBatch()
    p_far = beng.batchPairCount
    check_batch(beng, plain, [g], [3], far, "far cameras")
    # same sizes: only the view table changes; the recorded frame overflows and is rendered again
    beng.setViews(near)
    beng.rasterBatch()
    p_near = beng.batchPairCount
    say(f"[batch] pairs far {p_far} -> near {p_near}")
    assert p_near > 2 * p_far
    check_batch(beng, plain, [g], [3], near, "near cameras after growth")
    for _ in range(3):  # replays give the same frame
        beng.rasterBatch()
    check_batch(beng, plain, [g], [3], near, "near cameras replayed")
    # a new, smaller layout whose pairs fit the buffers already there: the frame's capacity is
    # restored from those buffers, not left at the zero a new layout starts from
    small = [view_tuple(sid, (0.0, 0.3, -2.5), 200, 120)]
    beng.setViews(small)
    beng.rasterBatch()
    check_batch(beng, plain, [g], [3], small, "smaller layout in the grown buffers")


def test_batch_entities_with_transforms(engines):
    from torpedo_amd.engine import Scene
    beng, plain = engines
    beng.clearScenes()
    a = small_scene(3000, seed=61)
    b = small_scene(2000, seed=62)
    sc = Scene()
    ea = sc.add(a)
    eb = sc.add(b)
    M = np.eye(4, dtype=np.float32)
    M[0, 3], M[1, 3] = 0.4, -0.3
    M[:3, :3] = np.array([[0.9, -0.2, 0.0], [0.2, 0.9, 0.1], [0.0, -0.1, 1.1]], np.float32)
    sid = beng.addScene(sc)
    beng.setSceneTransform(sid, eb, M)
    views = [view_tuple(sid, (0.2, 0.1, -3.0), 160, 120), view_tuple(sid, (-1.0, 0.5, -2.5), 200, 100)]
    beng.setViews(views)
    beng.rasterBatch()
    keys, vals = beng.readBatchPairs()
    plain.compile(sc)
    plain.getTransformHost().transform(eb, M)
    for vi, view in enumerate(views):
        tb, ib, _ = beng.batchViewLayout(vi)
        w, h = view[4], view[5]
        nt = ((w + 15) // 16) * ((h + 15) // 16)
        t = (keys >> np.uint64(32)).astype(np.int64)
        lo, hi = np.searchsorted(t, tb), np.searchsorted(t, tb + nt)
        pk, pv, pimg, _ = plain_frame(plain, view)
        lk = ((keys[lo:hi] >> np.uint64(32)) - np.uint64(tb)) << np.uint64(32) | (keys[lo:hi] & np.uint64(0xFFFFFFFF))
        assert np.array_equal(lk, pk) and np.array_equal(vals[lo:hi] - np.uint32(ib), pv), vi
        assert np.array_equal(beng.readBatchImage(vi), pimg), vi
    del ea


def test_batch_without_gaussians(engines):
    beng, _ = engines
    beng.clearScenes()
    sid = beng.addScene(np.zeros((0, 60), np.float32))
    views = [view_tuple(sid, (0, 0, -3), 48, 40), view_tuple(sid, (1, 0, -3), 17, 5)]
    beng.setViews(views)
    beng.rasterBatch()
    assert beng.batchPairCount == 0
    for vi in range(2):
        assert not beng.readBatchImage(vi).any()
        rgba = beng.readBatchRGBA8(vi)
        assert not rgba[..., :3].any() and np.all(rgba[..., 3] == 255)


def test_batch_all_items_culled(engines):
    """Cameras that look away from the scene: every item is culled, the item sort keeps none and the
    emission has no partition (its first block stores the zero totals); the frames before and after,
    which see the scene, still equal the plain frames."""
    from torpedo_amd import scenes
    from torpedo_amd.engine import PerspectiveCamera
    beng, plain = engines
    beng.clearScenes()
    g = scenes.cluster_scene(20_000, seed=11)
    sid = beng.addScene(g)
    sizes = [(160, 120), (96, 64), (33, 47)]
    eyes = scenes.fibonacci_sphere(len(sizes), 8.0)
    seen = [view_tuple(sid, eyes[i], w, h, up=(0.0, 0.0, 1.0)) for i, (w, h) in enumerate(sizes)]

    def away(i, w, h):  # at distance 30 on the -z axis, looking further away from the origin
        cam = PerspectiveCamera(w, h)
        cam.lookAt([0.0, 0.0, -30.0 - i], [0.0, 0.0, -60.0], [0.0, 1.0, 0.0])
        v, pv, fo = cam.uniforms()
        return (sid, v, pv, fo, w, h)

    hidden = [away(i, w, h) for i, (w, h) in enumerate(sizes)]
    try:
        for graph in (True, False):
            beng.setBatchGraph(graph)
            beng.setViews(seen)
            beng.rasterBatch()
            beng.syncBatch()
            check_batch(beng, plain, [g], [3], seen, f"seen before, graph={graph}")
            assert beng.batchPairCount > 0
            beng.setViews(hidden)  # same sizes: the layout and the recorded frame are kept
            beng.rasterBatch()
            beng.syncBatch()
            assert beng.batchPairCount == 0, graph
            assert beng.batchItemCounts() == (g.shape[0] * len(sizes), 0), graph
            for vi in range(len(sizes)):
                assert not beng.readBatchImage(vi).any(), (graph, vi)
            beng.setViews(seen)
            beng.rasterBatch()
            beng.syncBatch()
            check_batch(beng, plain, [g], [3], seen, f"seen after, graph={graph}")
    finally:
        beng.setBatchGraph(True)


def _plain_grads(plain, rec, deg, views, dls):
    """Sum over views of the plain backward for the scene rec."""
    n = rec.shape[0]
    plain.uploadRecords(rec, deg)
    bufs = [plain.alloc(max(n, 1) * k * 4) for k in (3, 3, 4, 1, 48)]
    d_dl = plain.alloc(max(max(d.nbytes for d in dls), 4))
    for k, (view, dl) in enumerate(zip(views, dls)):
        _, v, pv, fo, w, h = view
        plain.setCamera(v, pv, fo, w, h)
        plain.render()
        plain.upload(d_dl, dl)
        plain.backward(d_dl, *bufs, accumulate=k > 0)
    plain.waitIdle()
    out = []
    for b, k in zip(bufs, (3, 3, 4, 1, 48)):
        h = np.empty((n, k), np.float32)
        if n:
            plain.download(h, b)
        out.append(h)
        plain.free(b)
    plain.free(d_dl)
    return out


def _rel(a, b):
    d = np.linalg.norm(a.astype(np.float64) - b.astype(np.float64))
    s = np.linalg.norm(b.astype(np.float64))
    return d / s if s > 0 else d


@pytest.mark.parametrize("use_step", [False, True])
def test_batch_backward_equals_sum_of_plain(engines, use_step):
    beng, plain = engines
    beng.clearScenes()
    recs = [small_scene(2500, seed=71, log_scale_mean=-3.0), small_scene(1200, seed=72, log_scale_mean=-3.0)]
    degs = [3, 2]
    from torpedo_amd.engine import Settings
    ids = []
    for r, d in zip(recs, degs):
        s = Settings()
        s.sphericalHarmonicsDegree = d
        ids.append(beng.addScene(r, s))
    plan = [(0, (96, 64), (0.1, -0.2, -3.0)), (1, (80, 80), (0.0, 0.3, -2.8)), (0, (64, 100), (-0.6, 0.2, -2.9)),
            (0, (50, 40), (0.5, 0.5, -3.2))]
    views = [view_tuple(ids[s], eye, w, h) for s, (w, h), eye in plan]
    beng.setViews(views)
    rng = np.random.default_rng(5)
    dls = [rng.normal(size=(h, w, 3)).astype(np.float32) for _, (w, h), _ in plan]
    dl_all = np.concatenate([d.reshape(-1) for d in dls])
    d_dl = beng.alloc(dl_all.nbytes)
    beng.upload(d_dl, dl_all)
    grads = []
    host = []
    for r in recs:
        n = r.shape[0]
        bufs = [beng.alloc(n * k * 4) for k in (3, 3, 4, 1, 48)]
        grads.append(bufs)
    beng.setBatchGraph(True)
    for rep in range(2):
        if use_step:
            beng.batchStep(d_dl, grads, accumulate=rep > 0)
            beng.syncBatch()
        else:
            beng.rasterBatch(for_backward=True)
            beng.batchBackward(d_dl, grads, accumulate=rep > 0)
        beng.waitIdle()
    for s, r in enumerate(recs):
        n = r.shape[0]
        got = []
        for b, k in zip(grads[s], (3, 3, 4, 1, 48)):
            h = np.empty((n, k), np.float32)
            beng.download(h, b)
            got.append(h)
        sv = [i for i, (sc, _, _) in enumerate(plan) if sc == s]
        ref = _plain_grads(plain, r, degs[s], [views[i] for i in sv], [dls[i] for i in sv])
        for name, gg, rr in zip(("means", "log_scales", "quats", "opacity_logits", "sh"), got, ref):
            e = _rel(gg, 2.0 * rr)  # two accumulated steps (x + x = 2x exactly)
            say(f"[batch-bwd] step={use_step} scene {s} {name}: rel err {e:.2e}")
            assert same_bits(gg, 2.0 * rr), (s, name, e)
        host.append(got)
    for bufs in grads:
        for b in bufs:
            beng.free(b)
    beng.free(d_dl)


def test_two_frames_in_flight(engines):
    """Frame pipelining (SurfaceRenderer.h:66 keeps two frames in flight): two contexts on their own
    streams take frames in turn, each reading one device copy of the records, and a context is
    reused only after its frame from two steps back has been read.  Every frame must equal the
    plain frame of the same camera (FP32 and RGBA8 bit-exact)."""
    from torpedo_amd.engine import GaussianEngine
    _, plain = engines
    g = small_scene(20000, seed=31, log_scale_mean=-3.5)
    cams = []
    for i in range(7):
        a = 2.0 * np.pi * i / 7
        cams.append(view_tuple(0, [3.0 * np.sin(a), 0.4 * i - 1.2, -3.0 * np.cos(a)], 320 + 16 * i, 200 + 8 * i))
    engs = [GaussianEngine(), GaussianEngine()]
    d_rec = engs[0].alloc(g.nbytes)
    try:
        engs[0].upload(d_rec, g)
        for e in engs:
            e.clearScenes()
            e.addDeviceScene(d_rec, g.shape[0], 3)
            e.setBatchGraph(True)
        got = {}
        for i, c in enumerate(cams):
            e = engs[i % 2]
            if i >= 2:
                e.syncBatch()
                got[i - 2] = (e.readBatchImage(0), e.readBatchRGBA8(0))
            e.setViews([c])
            e.rasterBatch()
        for i in (len(cams) - 2, len(cams) - 1):
            e = engs[i % 2]
            e.syncBatch()
            got[i] = (e.readBatchImage(0), e.readBatchRGBA8(0))
        plain.uploadRecords(g, 3)
        for i, c in enumerate(cams):
            _, _, img, rgba = plain_frame(plain, c)
            say(f"[batch] in-flight frame {i}: {c[4]}x{c[5]}")
            assert np.array_equal(got[i][0].view(np.uint32), img.view(np.uint32)), i
            assert np.array_equal(got[i][1], rgba), i
    finally:
        engs[0].free(d_rec)
        for e in engs:
            e.close()


def test_batch_depths_outside_the_sort_depth_window(engines, orc):
    """A batch sorts its items on the depth-window key (depth bits - bits(2^-8), 27 bits) unless a
    depth of its layout lies outside [2^-8, 2^8), then on the 31-bit positive-float depth.  Checked
    three ways, each view bit-exact with the plain frame and the key shown by the pass count of a
    9-bit radix shape (3 item passes for the window key, 4 for the wider one, plus 1 pass over the
    260 tiles): a layout inside the window; a new layout whose first frame finds such depths; and a
    camera change under a kept graph that puts depths outside, which the frame itself detects (it
    is marked overflowed and rendered again on the wider key)."""
    from test_gpu_forward import depth_window_scene
    beng, plain = engines
    beng.clearScenes()
    recs = [small_scene(6000, seed=61, log_scale_mean=-3.0), depth_window_scene("far"), depth_window_scene("near")]
    sids = [beng.addScene(r) for r in recs]
    degs = [3, 3, 3]

    def view(sid, eye):
        cam = orc.make_camera(list(eye), [0.0, 0.0, 0.0], [0.0, -1.0, 0.0], 320, 200, near=0.001, far=1000.0)
        return (sid, cam.view, cam.proj_view, cam.focal_ndc, 320, 200)

    beng.setSortShape(6)  # 9-bit digits
    beng.setBatchGraph(True)
    beng.setProfiling(True)
    try:
        cases = [("inside", [view(sids[0], (0.0, 0.0, -3.0))], 4),
                 ("far depths, new layout", [view(sids[1], (0.0, 0.0, -3.0))], 5),
                 ("near depths, new layout", [view(sids[2], (0.0, 0.0, -3.0))], 5),
                 ("inside again", [view(sids[0], (0.0, 0.0, -3.0))], 4),
                 ("far camera, same layout", [view(sids[0], (0.0, 0.0, -400.0))], 5)]
        for name, views, want in cases:
            beng.setViews(views)
            beng.rasterBatch()
            beng.syncBatch()
            passes = beng.batchTimes()["radix_passes"]
            say(f"[batch] depth window, {name}: {passes} radix passes")
            check_batch(beng, plain, recs, degs, views, f"depth window, {name}")
            assert passes == want, (name, passes)
    finally:
        beng.setProfiling(False)


def _read_grads(eng, bufs, n):
    out = []
    for b, k in zip(bufs, (3, 3, 4, 1, 48)):
        h = np.empty((n, k), np.float32)
        eng.download(h, b)
        out.append(h)
    return out


def test_back_to_back_steps_after_an_overflow(engines):
    """Two batchStep calls issued without a synchronisation between them, the first of which
    overflows its pair buffers (the cameras moved closer under a kept layout): the second call must
    first render the first step again in grown buffers, so the accumulated gradients equal those of
    two synchronised steps."""
    beng, _ = engines
    beng.clearScenes()
    g = small_scene(60_000, seed=81, log_scale_mean=-3.0)
    n = g.shape[0]
    sid = beng.addScene(g)
    sizes = [(512, 512), (400, 300)]
    far = [view_tuple(sid, (0.0, 0.0, -30.0 - 5 * i), w, h) for i, (w, h) in enumerate(sizes)]
    near = [view_tuple(sid, (0.2 * i, 0.1, -2.6), w, h) for i, (w, h) in enumerate(sizes)]
    rng = np.random.default_rng(8)
    dl = np.concatenate([rng.normal(size=(h, w, 3)).astype(np.float32).reshape(-1) for w, h in sizes])
    d_dl = beng.alloc(dl.nbytes)
    beng.upload(d_dl, dl)
    bufs = [beng.alloc(n * k * 4) for k in (3, 3, 4, 1, 48)]
    try:
        beng.setBatchGraph(True)
        beng.setViews(far)
        beng.batchStep(d_dl, [bufs])
        beng.syncBatch()
        p_far = beng.batchPairCount
        beng.setViews(near)  # same sizes and scene: the layout and its pair capacity are kept
        beng.batchStep(d_dl, [bufs], accumulate=False)
        beng.batchStep(d_dl, [bufs], accumulate=True)  # no synchronisation in between
        beng.syncBatch()
        p_near = beng.batchPairCount
        say(f"[batch] back-to-back steps: pairs far {p_far} -> near {p_near}")
        # the far frame sized the pair buffers to (P + P/8) x 1.15 (at least 65536 x 1.15) pairs,
        # so the first near step overflowed them
        assert p_near > 1.3 * max(p_far, 65536), (p_far, p_near)
        got = _read_grads(beng, bufs, n)
        beng.batchStep(d_dl, [bufs], accumulate=False)
        beng.syncBatch()
        beng.batchStep(d_dl, [bufs], accumulate=True)
        beng.syncBatch()
        ref = _read_grads(beng, bufs, n)
        for name, a, b in zip(("means", "log_scales", "quats", "opacity_logits", "sh"), got, ref):
            e = _rel(a, b)
            say(f"[batch] back-to-back {name}: rel err {e:.2e}")
            assert np.abs(b).max() > 0 and same_bits(a, b), (name, e)
    finally:
        for b in bufs:
            beng.free(b)
        beng.free(d_dl)
