"""Forward path on the GPU (through the C ABI) against the CPU oracle.

Bar: splats, tile counts, offsets, keys, values and ranges bit-exact; the FP32 image bit-exact
with the oracle (the test asserts max-abs <= 1e-5 and reports the number of differing pixels);
RGBA8 exact.  Every image's error against the FP64 twin is stated (printed, and logged when
TPD_FP64_LOG names a file); its 99.9th percentile must stay below 1e-4.
"""
import sys

import numpy as np
import pytest

from conftest import c1_camera, report_fp64, small_scene

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    from torpedo_amd.engine import GaussianEngine
    e = GaussianEngine(keep_unsorted=True)
    yield e
    e.close()


def gpu_render(eng, g, cam, deg):
    eng.uploadRecords(g, deg)
    eng.setCamera(cam.view, cam.proj_view, cam.focal_ndc, cam.width, cam.height)
    eng.render()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def compare_frame(eng, orc, g, cam, deg, name, check_f64=True):
    say(f"[forward] {name}: N={g.shape[0]} {cam.width}x{cam.height} deg={deg}")
    gpu_render(eng, g, cam, deg)
    ref = orc.render(g, cam, deg)
    sp = eng.readSplats()
    for f in ("color", "texel", "copac"):
        diff = np.nonzero(np.any(bits(sp[f]) != bits(ref["splats"][f]), axis=1))[0]
        assert diff.size == 0, (name, f, diff[:5], sp[f][diff[:3]], ref["splats"][f][diff[:3]])
    assert np.array_equal(sp["tiles"], ref["splats"]["tiles"]), name
    counts, offs = eng.readTiles()
    assert np.array_equal(offs, ref["offsets"]), name
    assert eng.pairCount == ref["total"], name
    uk, uv = eng.readPairs(unsorted=True)
    assert np.array_equal(uk, ref["keys_unsorted"]) and np.array_equal(uv, ref["vals_unsorted"]), name
    k, v = eng.readPairs()
    assert np.array_equal(k, ref["keys"]) and np.array_equal(v, ref["vals"]), name
    assert np.array_equal(eng.readRanges(), ref["ranges"]), name
    img = eng.readImage()
    err = np.abs(img - ref["image"])
    ndiff = int(np.count_nonzero(err))
    say(f"[forward] {name}: P={ref['total']} image max-abs vs oracle {err.max():.3g} ({ndiff} values differ)")
    assert err.max() <= IMG_TOL, name
    assert np.array_equal(eng.readRGBA8(), ref["rgba8"]), name
    if check_f64:
        _, img64 = orc.render_f64(g, cam, deg)
        assert report_fp64(name, img, img64)[2] < 1e-4, name
    return ref


def test_plain_forward_100k_720p(eng, orc):
    from torpedo_amd import scenes
    g = scenes.uniform_scene(100_000, seed=1)
    compare_frame(eng, orc, g, c1_camera(orc), 3, "uniform-100k-1280x720")


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_degrees(eng, orc, deg):
    g = small_scene(4000, seed=20 + deg)
    compare_frame(eng, orc, g, c1_camera(orc, 320, 180), deg, f"sh-degree-{deg}")


@pytest.mark.parametrize("w,h", [(1, 1), (17, 33), (64, 64), (250, 130)])
def test_ragged_image_sizes(eng, orc, w, h):
    g = small_scene(3000, seed=w * 7 + h, log_scale_mean=-3.5)
    compare_frame(eng, orc, g, c1_camera(orc, w, h), 3, f"size-{w}x{h}")


def test_edge_scenes(eng, orc):
    from torpedo_amd.engine import gaussian_point
    cam = c1_camera(orc, 160, 96)
    # a single big Gaussian covering every tile
    big = gaussian_point([0, 0, 0], 0.9, (0, 0, 0, 1), (2.0, 2.0, 2.0), np.full(48, 0.1, np.float32))[None]
    compare_frame(eng, orc, big, cam, 3, "one-giant-gaussian")
    # everything culled (behind the camera)
    g = small_scene(1000, seed=3)
    g[:, 2] -= 10.0
    ref = compare_frame(eng, orc, g, cam, 3, "all-culled")
    assert ref["total"] == 0 and np.all(eng.readImage() == 0)
    # empty scene renders black (the reference warns and renders nothing)
    say("[forward] empty-scene")
    gpu_render(eng, np.zeros((0, 60), np.float32), cam, 3)
    assert eng.pairCount == 0 and np.all(eng.readImage() == 0)
    assert np.all(eng.readRGBA8()[..., 3] == 255) and np.all(eng.readRGBA8()[..., :3] == 0)
    assert np.all(eng.readRanges() == 0)
    # opacity 0 and 1, duplicates with identical depth (stable order by index)
    g = small_scene(2000, seed=4)
    g[:500, 3] = 0.0
    g[500:1000, 3] = 1.0
    g[1000:1500] = g[1500:2000]
    compare_frame(eng, orc, g, cam, 3, "opacity-extremes-and-duplicates")


def test_nearly_equal_depths_every_sort_shape(eng, orc):
    """Keys that agree in all but their lowest bits: view depths 3 + k ulp(3) (k < 256) put
    thousands of pairs with nearly equal keys in the centre tiles and a few colliding ones in
    every other tile.  Every sort shape (8- and 9-bit digits, every block shape) must give the
    oracle's stable order (tile, depth, emission)."""
    from torpedo_amd.engine import gaussian_point
    cam = c1_camera(orc, 160, 96)  # view = [I | (0, 0, 3)]: view z = world z + 3, exactly
    rng = np.random.default_rng(11)
    ulp = float(np.spacing(np.float32(3.0)))
    sh = np.full(48, 0.2, np.float32)
    recs = []
    for _ in range(3000):  # around the view axis: thousands of pairs in the centre tiles
        x, y = rng.normal(0.0, 0.05, 2)
        recs.append(gaussian_point([x, y, float(rng.integers(0, 256)) * ulp], 0.3, (0, 0, 0, 1),
                                   (0.01, 0.01, 0.01), sh))
    for _ in range(600):  # scattered: a few pairs per tile with colliding depths
        x, y = rng.uniform(-0.8, 0.8, 2)
        recs.append(gaussian_point([x, y * 0.6, float(rng.integers(0, 6)) * ulp], 0.3, (0, 0, 0, 1),
                                   (0.01, 0.01, 0.01), sh))
    g = np.stack(recs).astype(np.float32)
    from test_gpu_sort import NUM_SHAPES
    try:
        for shape in range(NUM_SHAPES):
            eng.setSortShape(shape)
            compare_frame(eng, orc, g, cam, 3, f"equal-depth-runs shape {shape}")
    finally:
        eng.setSortShape(0)


def test_buffer_growth_replay(eng, orc):
    cam = c1_camera(orc, 320, 240)
    compare_frame(eng, orc, small_scene(500, seed=8), cam, 3, "growth-small")
    compare_frame(eng, orc, small_scene(60000, seed=9, log_scale_mean=-2.5), cam, 3, "growth-large")


def test_entities_and_transforms(eng, orc):
    from torpedo_amd.engine import PerspectiveCamera, Scene, Settings, gaussian_point
    a = small_scene(3000, seed=30)
    b = small_scene(2000, seed=31)
    single = gaussian_point([0.2, 0.1, 0.0], 0.8, (0, 0, 0, 1), (0.2, 0.1, 0.3), np.full(48, 0.05, np.float32))
    scene = Scene()
    ea = scene.add(a)
    eb = scene.add(b)
    es = scene.add(single)
    say("[forward] entities-and-transforms")
    eng.compile(scene, Settings(3))
    T1 = np.array([[1, 0, 0, 0.3], [0, 1, 0, -0.2], [0, 0, 1, 0.1], [0, 0, 0, 1]], np.float32)
    c, s = np.cos(0.3), np.sin(0.3)
    T2 = np.array([[c, 0, s, 0], [0, 1, 0, 0.1], [-s, 0, c, 0], [0, 0, 0, 1]], np.float32)
    eng.getTransformHost().transform(eb, T1)
    eng.getTransformHost().transform(es, T2)
    eng.getTransformHost().transform(12345, T2)  # unknown entity: ignored
    cam = PerspectiveCamera(320, 180)
    cam.lookAt([0.3, -0.5, -3.0], [0, 0, 0], [0, -1, 0])
    eng.rasterFrame(cam)
    view, pv, fo = cam.uniforms()
    ocam = orc.Camera(view, pv, fo, 320, 180)
    emap = scene.buildEntityMap()
    assert emap == {eb: 0, ea: 1, es: 2}  # the last added group first, then the single point
    models = np.zeros((3, 16), np.float32)
    models[emap[ea]] = np.eye(4, dtype=np.float32).reshape(16)
    models[emap[eb]] = T1.reshape(16)
    models[emap[es]] = T2.reshape(16)
    ref = orc.render(scene.dataAll(), ocam, 3, ent_idx=scene.entityIndices(), models=models)
    k, v = eng.readPairs()
    assert np.array_equal(k, ref["keys"]) and np.array_equal(v, ref["vals"])
    assert np.abs(eng.readImage() - ref["image"]).max() <= IMG_TOL
    assert ea == 0


def test_bind_device_scene_after_an_entity_upload(eng, orc):
    """A scene bound from device records is one entity with the identity model, whatever an earlier
    compile left: here a one-entity scene of n points with a non-identity transform, then a bound
    scene of 2n records (the earlier per-point entity indices cover only n of them)."""
    from torpedo_amd._capi import check, lib, ptr
    a = small_scene(3000, seed=32)
    say("[forward] bind after entity upload")
    M = np.eye(4, dtype=np.float32)
    M[0, 3] = 0.7
    idx = np.zeros(a.shape[0], np.uint32)  # as the C++ compile uploads a one-entity scene
    check(lib().tpd_upload(eng._ctx, ptr(a), a.shape[0], ptr(idx), ptr(np.eye(4, dtype=np.float32).reshape(16)), 1, 3))
    check(lib().tpd_set_transform(eng._ctx, 0, ptr(M.reshape(16))))
    g = small_scene(6000, seed=33)
    d = eng.alloc(g.nbytes)
    try:
        eng.upload(d, g)
        eng.bindDeviceScene(d, g.shape[0], 3)
        cam = c1_camera(orc, 320, 180)
        eng.setCamera(cam.view, cam.proj_view, cam.focal_ndc, cam.width, cam.height)
        eng.render()
        ref = orc.render(g, cam, 3)
        k, v = eng.readPairs()
        assert np.array_equal(k, ref["keys"]) and np.array_equal(v, ref["vals"])
        assert np.abs(eng.readImage() - ref["image"]).max() <= IMG_TOL
    finally:
        eng.uploadRecords(np.zeros((0, 60), np.float32), 3)  # unbind before the records are freed
        eng.free(d)


def test_cluster_scene_1m_1080p(eng, orc):
    from torpedo_amd import scenes
    g = scenes.cluster_scene(1_000_000, seed=2)
    cam = orc.make_camera([0, 0, -8], [0, 0, 0], [0, -1, 0], 1920, 1080)
    ref = compare_frame(eng, orc, g, cam, 3, "clusters-1M-1920x1080")
    k = ref["keys"]
    assert np.all(k[1:] >= k[:-1])


@pytest.mark.parametrize("w,h", [(320, 200), (1920, 1080), (4096, 2304)])
def test_tile_order_is_a_heaviest_first_permutation(eng, orc, w, h):
    """The blend's tile launch order (1, 8 and 36 chunks of 1024 tiles): every tile exactly once,
    and the bit length of the tile's pair count never increases along the order."""
    g = small_scene(30000, seed=51, log_scale_mean=-3.5)
    cam = c1_camera(orc, w, h)
    say(f"[forward] tile order {w}x{h}")
    gpu_render(eng, g, cam, 3)
    order = eng.readTileOrder().astype(np.int64)
    r = eng.readRanges().astype(np.int64)
    n = r[:, 1] - r[:, 0]
    assert np.array_equal(np.sort(order), np.arange(n.size))
    cls = np.where(n > 0, np.floor(np.log2(np.maximum(n, 1))).astype(np.int64) + 1, 0)
    assert np.all(np.diff(cls[order]) <= 0)
    assert n.max() > 0 and cls[order[0]] == cls.max()


def depth_window_scene(case):
    """4000 Gaussians around the origin plus 2000 at view depths outside [2^-8, 2^8) for a camera at
    (0, 0, -3): behind the origin at depth ~400 ("far") or just in front of the eye at depth ~0.003
    ("near"); scales grown / shrunk with the distance so that they still cover pixels."""
    g = small_scene(6000, seed=61, log_scale_mean=-3.0)
    out = g[4000:]
    if case == "far":
        out[:, 0:3] = out[:, 0:3] * 40.0 + np.array([0.0, 0.0, 400.0], np.float32)
        out[:, 8:11] *= 130.0
    else:
        out[:, 0:3] = out[:, 0:3] * 0.0006 + np.array([0.0, 0.0, -3.0 + 0.003], np.float32)
        out[:, 8:11] *= 0.001
    return g


def depth_edge_scene(depths):
    """small_scene(6000) plus one Gaussian at each given view depth on the axis of a camera at
    (0, 0, -3) (view z = world z + 3 exactly), sized to cover a few pixels."""
    from torpedo_amd.engine import gaussian_point
    g = small_scene(6000, seed=61, log_scale_mean=-3.0)
    extra = [gaussian_point([0.0, 0.0, float(np.float32(d) - np.float32(3.0))], 0.8, (0, 0, 0, 1),
                            (0.004 * d,) * 3, np.full(48, 0.3, np.float32)) for d in depths]
    for d, r in zip(depths, extra):
        assert np.float32(r[2]) + np.float32(3.0) == np.float32(d)  # the view depth is exactly d
    return np.concatenate([g, np.stack(extra).astype(np.float32)])


@pytest.mark.parametrize("case", ["far", "near", "inside", "edges-inside", "edge-256", "edge-below"])
def test_depths_outside_the_sort_depth_window(eng, orc, case):
    """The sort keys on (tile << 27) | (depth bits - bits(2^-8)) when every emitted depth lies in
    [2^-8, 2^8), else on (tile << 31) | depth bits.  Frames with depths beyond either end must take
    the wider key and stay bit-exact with the oracle; with a 9-bit radix shape the pass count shows
    which key ran (320x200: 260 tiles = 9 bits; 9 + 27 bits in 4 passes, 9 + 31 in 5)."""
    edge = 2.0 ** -8
    if case in ("far", "near"):
        g = depth_window_scene(case)
    elif case == "inside":
        g = small_scene(6000, seed=61, log_scale_mean=-3.0)
    elif case == "edges-inside":  # the window's first and last float: [2^-8, 2^8 - 2^-16]
        g = depth_edge_scene([edge, 256.0 - 2.0 ** -16])
    elif case == "edge-256":  # the first float past the window
        g = depth_edge_scene([256.0])
    else:  # just below it: 2^-8 - 2^-22, the closest depth to 2^-8 that world z = d - 3 reaches
        g = depth_edge_scene([edge - 2.0 ** -22])
    cam = orc.make_camera([0.0, 0.0, -3.0], [0.0, 0.0, 0.0], [0.0, -1.0, 0.0], 320, 200, near=0.001, far=1000.0)
    sp = orc.render(g, cam, 3)["splats"]
    ref_depths, vis = sp["texel"][:, 2], sp["tiles"] > 0
    inside_case = case in ("inside", "edges-inside")
    outside = (ref_depths[vis] < 2.0 ** -8) | (ref_depths[vis] >= 2.0 ** 8)
    assert outside.any() == (not inside_case), (case, ref_depths[vis].min(), ref_depths[vis].max())
    if case == "edges-inside":  # both edge Gaussians are visible and keep their exact depths
        assert np.any(ref_depths[vis] == np.float32(2.0 ** -8)) and np.any(ref_depths[vis] == np.float32(256.0 - 2.0 ** -16))
    eng.setSortShape(6)  # 9-bit digits, 512 x 12
    eng.setProfiling(True)
    try:
        compare_frame(eng, orc, g, cam, 3, f"depth window {case}")
        passes = eng.stageTimes()["radix_passes"]
    finally:
        eng.setProfiling(False)
    assert passes == (4 if inside_case else 5), (case, passes)
