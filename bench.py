"""Benchmark of the Gaussian rasterization path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one batched forward frame of the workload "multiview-32-clusters-1M": one scene of 1M
Gaussians in 64 anisotropic clusters (seeded, synthetic) seen by 32 cameras on a Fibonacci sphere
of radius 8 looking at the origin, resolutions drawn (seed 4) from 256x256 ... 1920x1080, SH
degree 3, 16x16 tiles, all 32 views in one replay of the recorded HIP graph (launch + wait).
`value` is views per second of that batched frame; the same 32 views rendered as 32 sequential
plain frames (render + wait each) give the plain rate beside it.

Also measured (each printed in the same JSON line):
  plain_1080p            the same scene from one camera at distance 8, 1920x1080 (plain frames)
  sort-load variant      log-scale mean -3 (~18M pairs) for the sort's HBM fraction
  fwd_bwd_3M_1080p       forward + backward of 3M clustered Gaussians at 1920x1080, dL/dpixel ~ N(0,1)
  multiscene_16x4_step   16 scenes (log-uniform 50k..2M Gaussians) x 4 views (512..1600 x 512..900),
                         forward + backward of all 64 views in one recorded graph
Stage times come from HIP events recorded on the render stream.

One JSON line is printed.  Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import subprocess
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "views/s batched vs plain; fwd+bwd ms @3M Gaussians 1080p; sort % HBM peak"
HBM_PEAK_GBS = 8000.0
# HBM bytes per element per k_onesweep launch of the batched frame (32-bit keys and values; mean
# over the frame's item and pair passes), per radix shape, from two rocprofv3 --pmc runs:
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors read off a plain copy kernel).
# HBM bytes per k_onesweep launch of the batched frame (mean over its launches), measured with
# FETCH_SIZE x 2 + WRITE_SIZE for the (item sort, pair sort) block shapes named; other shapes: null.
ONESWEEP_TRAFFIC = {
    (10, 9): {"bytes_per_launch": 779_200_000,
              "source": "profiles/batch_frame-multiview32-clusters1M.md (item passes: first 614.2 MB, later "
                        "2 x 626.1 MB; pair passes 2 x 1,014.9 MB; FETCH_SIZE x 2 + WRITE_SIZE)"},
    (6, 8): {"bytes_per_launch": 893_200_000,
             "source": "profiles/batch_frame-multiview32-clusters1M.md (item passes: first 731.9 MB, later "
                       "2 x 708.0 MB; pair passes 2 x 1,159.1 MB; FETCH_SIZE x 2 + WRITE_SIZE)"},
}
VALU_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def camera_uniforms(eye, w, h, up=(0.0, -1.0, 0.0)):
    from torpedo_amd.engine import PerspectiveCamera
    cam = PerspectiveCamera(w, h)
    cam.lookAt(list(eye), [0, 0, 0], list(up))
    return cam.uniforms()


def med(xs):
    return float(np.median(np.asarray(xs, np.float64)))


def multiview_views(scene_id=0):
    from torpedo_amd import scenes
    out = []
    for eye, (w, h) in scenes.multiview_rig(32, 8.0, seed=4):
        v, pv, fo = camera_uniforms(eye, w, h, up=(0.0, 0.0, 1.0))
        out.append((scene_id, v, pv, fo, w, h))
    return out


# Radix-pass block shapes in the library's table order: (digit bits, threads, keys per thread).
SORT_SHAPES = [(8, 256, 16), (8, 256, 8), (8, 256, 12), (8, 512, 8), (8, 512, 12), (8, 256, 20), (9, 512, 12),
               (9, 512, 8), (9, 512, 16), (9, 1024, 16), (9, 1024, 12), (8, 1024, 12)]


def shape_desc(i):
    if i is None or i < 0 or i >= len(SORT_SHAPES):
        return None
    b, t, k = SORT_SHAPES[i]
    return {"index": i, "digit_bits": b, "threads": t, "keys_per_thread": k}


def sort_keys_per_s(P, stage_ms):
    """Pairs sorted per second over the whole sort: histogram pass plus every radix pass."""
    t = (stage_ms["sort_histogram"] + stage_ms["sort_passes"]) / 1000.0
    return P / t if t > 0 else 0.0


def onesweep_roofline(P, sort_passes_ms, passes):
    t_launch = sort_passes_ms / 1000.0 / max(passes, 1)
    achieved = 24.0 * P / t_launch / 1e9 if t_launch > 0 else 0.0
    return achieved, t_launch


def time_multiview(eng, g, steps, warmup, tune=True):
    """32 views of one scene: 32 sequential plain frames vs one batched frame."""
    views = multiview_views()
    eng.uploadRecords(g, 3)
    eng.setProfiling(False)

    def plain_pass():
        t0 = time.perf_counter()
        for (_, v, pv, fo, w, h) in views:
            eng.setCamera(v, pv, fo, w, h)
            eng.render()
            eng.waitIdle()
        return time.perf_counter() - t0

    plain_tuned = None
    if tune:  # the plain frames get their best radix shape too, tuned on the largest view
        (_, v, pv, fo, w, h) = max(views, key=lambda x: x[4] * x[5])
        eng.setCamera(v, pv, fo, w, h)
        eng.render()
        plain_tuned = eng.autotune()
    for _ in range(max(warmup, 1)):
        plain_pass()
    t_plain = med([plain_pass() for _ in range(max(3, steps // 4))])
    plain_pairs = 0
    for (_, v, pv, fo, w, h) in views:
        eng.setCamera(v, pv, fo, w, h)
        eng.render()
        plain_pairs += eng.pairCount

    eng.clearScenes()
    eng.addScene(g)
    eng.setViews(views)
    eng.setBatchGraph(True)
    eng.rasterBatch()
    eng.syncBatch()
    tuned = eng.autotuneBatch() if tune else None
    for _ in range(warmup):
        eng.rasterBatch()
        eng.syncBatch()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.rasterBatch()
        eng.syncBatch()
    t_batch = (time.perf_counter() - t0) / steps
    P = eng.batchPairCount
    eng.setProfiling(True)
    gpu = []
    for _ in range(5):
        eng.rasterBatch()
        eng.syncBatch()
        gpu.append(eng.batchTimes()["frame"])
    eng.setBatchGraph(False)
    st, so = [], []
    for _ in range(3):
        eng.rasterBatch()
        eng.syncBatch()
        st.append(eng.batchTimes())
        so.append(eng.batchSortTimes())
    evals = eng.batchBlendEvaluations()  # counted by one more blend of the last frame, not timed
    eng.setBatchGraph(True)
    eng.setProfiling(False)
    stages = {k: round(med([s[k] for s in st]), 4) for k in st[0]}
    sorts = {k: (round(med([s[k] for s in so]), 4) if k.endswith("_ms") else so[0][k]) for k in so[0]}
    sorts["items_kept"] = eng.batchItemCounts()[1]  # the visible items: the item sort drops the rest
    return {"views": len(views), "batched_views_per_s": round(len(views) / t_batch, 2),
            "plain_views_per_s": round(len(views) / t_plain, 2),
            "batched_over_plain": round(t_plain / t_batch, 3),
            "batched_ms_per_frame_wall": round(t_batch * 1000.0, 4), "batched_ms_gpu": round(med(gpu), 4),
            "plain_ms_32_views_wall": round(t_plain * 1000.0, 4), "pairs": P, "plain_pairs_sum": plain_pairs,
            "eager_stage_ms": stages, "sorts": sorts, "blend_evaluations": evals, "autotune": tuned,
            "autotune_plain": plain_tuned,
            "resolutions": [[v[4], v[5]] for v in views]}


def time_plain(eng, g, eye, w, h, steps, warmup, tune=False):
    eng.uploadRecords(g, 3)
    view, pv, fo = camera_uniforms(eye, w, h)
    eng.setCamera(view, pv, fo, w, h)
    eng.setProfiling(True)
    tuned = None
    if tune:
        eng.render()
        tuned = eng.autotune()
    for _ in range(warmup):
        eng.render()
    eng.waitIdle()
    stages = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.render()
        stages.append(eng.stageTimes())  # waits for the frame's last event
    eng.waitIdle()
    dt = time.perf_counter() - t0
    P = eng.pairCount
    out = {k: med([s[k] for s in stages]) for k in stages[0]}
    out["blend_evaluations"] = eng.blendEvaluations()  # the last frame, blended once more, not timed
    if tuned is not None:
        out["autotune"] = tuned
    return dt, P, out


def time_fwd_bwd(eng, steps, warmup):
    """Forward + backward of 3M clustered Gaussians at 1920x1080 with dL/dpixel ~ N(0,1): the span
    from the first forward kernel to the end of the parameter-gradient pass (HIP events)."""
    from torpedo_amd import scenes
    W, H = 1920, 1080
    g = scenes.cluster_scene(3_000_000, seed=3)
    n = g.shape[0]
    eng.uploadRecords(g, 3)
    view, pv, fo = camera_uniforms([0.0, 0.0, -8.0], W, H)
    eng.setCamera(view, pv, fo, W, H)
    eng.render()
    tuned = eng.autotune()  # the radix shape for this frame's keys (not timed)
    eng.setProfiling(True)
    up = np.random.default_rng(3).standard_normal((H, W, 3)).astype(np.float32)
    d_up = eng.alloc(up.nbytes)
    eng.upload(d_up, up)
    bufs = [eng.alloc(n * k * 4) for k in (3, 3, 4, 1, 48)]
    for _ in range(warmup):
        eng.render()
        eng.backward(d_up, *bufs)
    eng.waitIdle()
    rec = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.render()
        eng.backward(d_up, *bufs)
        rec.append({**eng.stageTimes(), **eng.backwardTimes()})
    eng.waitIdle()
    wall = (time.perf_counter() - t0) / steps * 1000.0
    P = eng.pairCount
    for b in bufs + [d_up]:
        eng.free(b)
    eng.setProfiling(False)
    m = {k: round(med([r[k] for r in rec]), 4) for k in rec[0]}
    return {"ms": m["fwd_bwd_span"], "wall_ms_per_step": round(wall, 4), "pairs": P, "gaussians": n,
            "width": W, "height": H, "stage_ms": {k: v for k, v in m.items() if k not in ("radix_passes",)},
            "autotune": tuned}


def pipeline_cameras(n, w=1920, h=1080):
    """n cameras on a circle of radius 8 around the clusters-1M scene (one frame each)."""
    out = []
    for i in range(n):
        a = 2.0 * np.pi * i / n
        v, pv, fo = camera_uniforms([8.0 * np.sin(a), 0.0, -8.0 * np.cos(a)], w, h)
        out.append((0, v, pv, fo, w, h))
    return out


def time_pipelined(g, frames, warmup):
    """Frame pipelining: a stream of single-view 1920x1080 frames of the clusters-1M scene, (a) one context,
    each frame set up, launched and waited for, (b) two contexts on their own streams taking frames
    in turn, so that frame i + 1 is launched while frame i still runs and the host waits only for
    frame i - 1 before reusing its context.  Both contexts read one device copy of the records."""
    from torpedo_amd.engine import GaussianEngine
    cams = pipeline_cameras(frames + warmup)
    engs = [GaussianEngine(device=0), GaussianEngine(device=0)]
    d_rec = engs[0].alloc(g.nbytes)
    engs[0].upload(d_rec, g)
    for e in engs:
        e.clearScenes()
        e.addDeviceScene(d_rec, g.shape[0], 3)
        e.setBatchGraph(True)
        e.setViews([cams[0]])
        e.rasterBatch()
        e.syncBatch()
    # (a) one context, one frame at a time
    a = engs[0]
    for c in cams[:warmup]:
        a.setViews([c]); a.rasterBatch(); a.syncBatch()
    t0 = time.perf_counter()
    for c in cams[warmup:]:
        a.setViews([c]); a.rasterBatch(); a.syncBatch()
    t_seq = (time.perf_counter() - t0) / frames
    # (b) two contexts, two frames in flight
    def run(seq):
        for i, c in enumerate(seq):
            e = engs[i % 2]
            e.setViews([c])  # waits for this context's previous frame (two frames back)
            e.rasterBatch()
        for e in engs:
            e.syncBatch()
    run(cams[:warmup])
    t0 = time.perf_counter()
    run(cams[warmup:])
    t_pipe = (time.perf_counter() - t0) / frames
    pairs = engs[0].batchPairCount
    engs[0].free(d_rec)
    for e in engs:
        e.close()
    return {"frames": frames, "width": 1920, "height": 1080, "pairs_last_frame": pairs,
            "sequential_ms_per_frame": round(t_seq * 1000.0, 4), "sequential_frames_per_s": round(1.0 / t_seq, 2),
            "pipelined_ms_per_frame": round(t_pipe * 1000.0, 4), "pipelined_frames_per_s": round(1.0 / t_pipe, 2),
            "pipelined_over_sequential": round(t_seq / t_pipe, 3)}


def time_multiscene(eng, steps, warmup):
    """16 scenes x 4 views, forward + backward of all 64 views in one recorded graph."""
    from torpedo_amd import scenes
    sizes = scenes.multiscene_sizes(16, 5)
    rigs = scenes.multiscene_views(16, 4, 8.0, 5)
    eng.clearScenes()
    for s, n in enumerate(sizes):
        eng.addScene(scenes.cluster_scene(n, seed=5 + s))
    views = []
    for s in range(16):
        for eye, (w, h) in rigs[s]:
            v, pv, fo = camera_uniforms(eye, w, h, up=(0.0, 0.0, 1.0))
            views.append((s, v, pv, fo, w, h))
    eng.setViews(views)
    tuned = eng.autotuneBatch()  # both sorts' block shapes for this layout (forward frames, not timed)
    npix = sum(v[4] * v[5] for v in views)
    dl = np.random.default_rng(5).standard_normal(3 * npix).astype(np.float32)
    d_dl = eng.alloc(dl.nbytes)
    eng.upload(d_dl, dl)
    grads = [[eng.alloc(max(n, 1) * k * 4) for k in (3, 3, 4, 1, 48)] for n in sizes]
    eng.setBatchGraph(True)
    eng.setProfiling(False)
    for _ in range(max(warmup, 1)):
        eng.batchStep(d_dl, grads)
        eng.syncBatch()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.batchStep(d_dl, grads)
        eng.syncBatch()
    wall = (time.perf_counter() - t0) / steps
    P = eng.batchPairCount
    eng.setProfiling(True)
    gpu = []
    for _ in range(3):
        eng.batchStep(d_dl, grads)
        eng.syncBatch()
        gpu.append(eng.batchTimes()["frame"])
    eng.setProfiling(False)
    for gs in grads:
        for b in gs:
            eng.free(b)
    eng.free(d_dl)
    eng.clearScenes()
    return {"scenes": 16, "views": len(views), "gaussians": int(sum(sizes)), "pixels": int(npix), "pairs": P,
            "ms_per_step_wall": round(wall * 1000.0, 4), "ms_per_step_gpu": round(med(gpu), 4),
            "views_per_s": round(len(views) / wall, 2), "autotune": tuned}


def hbm_ceiling(eng):
    """Achievable HBM bandwidth on this GPU: copy / hipMemcpyAsync / nontemporal copy / read-only /
    write-only streams of 256 MiB .. 4 GiB, kernels at 2048 .. 32768 blocks of 256 threads (GB/s,
    copies counting read + written bytes); and the rate of random 8- and 4-byte reads from arrays
    of 32 MiB .. 2 GiB (G reads/s)."""
    out = {}
    for nbytes in (256 << 20, 1 << 30, 4 << 30):
        for kind in eng.BANDWIDTH_KINDS:
            for blocks in ((0,) if kind == "memcpy" else (2048, 8192, 32768)):
                key = f"{kind}/{nbytes >> 20}MiB" + (f"/{blocks}blocks" if blocks else "")
                out[key] = round(eng.measureBandwidth(kind, nbytes, 20, blocks), 1)
                log(f"[bench] {key}: {out[key]} GB/s")
    gathers = {}
    for nbytes in (32 << 20, 128 << 20, 256 << 20, 512 << 20, 2 << 30):
        for kind in eng.GATHER_KINDS:
            key = f"{kind}/{nbytes >> 20}MiB"
            gathers[key] = round(eng.measureGather(kind, nbytes), 2)
            log(f"[bench] {key}: {gathers[key]} G reads/s")
    return {"hbm_ceiling_gbs": out, "random_reads_g_per_s": gathers, "peak_gbs": HBM_PEAK_GBS}


def run_scatter(eng):
    """The memory pattern of a radix pass without its ranking (tpd_measure_run_scatter), at the
    sizes and block shapes of the sorts it bounds: keys and values read coalesced and written into
    per-(digit, block) runs with uniformly random digits.  Each case gives GB/s (bytes read +
    written) and the time per pass; the pair case also with 1, 2 and 4 resident blocks per CU."""
    cases = [  # (name, key bytes, keys, digit bins, keys per block, blocks per CU to try)
        ("batch pair pass (58.3M, 32-bit keys, 9-bit digits, 1024 x 16)", 4, 58_288_090, 512, 16384, (0, 1, 2, 4)),
        ("batch pair pass (58.3M, 32-bit keys, 9-bit digits, 512 x 16)", 4, 58_288_090, 512, 8192, (0,)),
        ("batch item pass (24.2M kept, 64-bit keys, 9-bit digits, 1024 x 12)", 8, 24_166_852, 512, 12288, (0,)),
        ("batch item pass (24.2M kept, 64-bit keys, 9-bit digits, 512 x 12)", 8, 24_166_852, 512, 6144, (0,)),
        ("plain 1080p 32-bit pass (21.1M, 8-bit digits, 1024 x 12)", 4, 21_122_859, 256, 12288, (0,)),
        ("plain 1080p 64-bit pass (21.1M, 8-bit digits, 1024 x 12)", 8, 21_122_859, 256, 12288, (0,)),
        ("plain 1080p 32-bit pass (21.1M, 8-bit digits, 512 x 12)", 4, 21_122_859, 256, 6144, (0,)),
    ]
    for kpb in (4096, 6144, 12288):  # run length 8, 12, 24 keys per (digit, block); 16 and 32 above
        cases.append((f"pair-sized sweep (58.3M, 32-bit keys, 9-bit digits, {kpb} keys per block)", 4, 58_288_090,
                      512, kpb, (0,)))
    out = []
    for name, kb, n, bins, kpb, caps in cases:
        n = n // kpb * kpb
        for bpc in caps:
            gbs = eng.measureRunScatter(kb, n, bins, kpb, bpc)
            us = 2.0 * n * (kb + 4) / (gbs * 1e9) * 1e6
            out.append({"case": name, "keys": n, "key_bytes": kb, "bins": bins, "keys_per_block": kpb,
                        "keys_per_run": kpb / bins, "blocks_per_cu": bpc, "gbs": round(gbs, 1),
                        "us_per_pass": round(us, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 3)})
            log(f"[bench] run scatter {name}, {bpc or 'any'} blocks/CU: {out[-1]['gbs']} GB/s, "
                f"{out[-1]['us_per_pass']} us")
    return {"run_scatter": out, "peak_gbs": HBM_PEAK_GBS}


def sort_vs_rocprim():
    """tests/cpp/sort_vs_rocprim (built by __graft_entry__.build()): the stand-alone sort under every
    block shape against rocPRIM's radix_sort_pairs, same keys, values and key bits."""
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "cpp", "sort_vs_rocprim")
    if not os.path.exists(exe):
        raise SystemExit("tests/cpp/sort_vs_rocprim is not built (run __graft_entry__.build())")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise SystemExit(f"sort_vs_rocprim failed ({r.returncode}): {r.stderr.strip()}")
    cases = [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]
    for c in cases:
        best = min(range(len(c["tpd_ms_per_shape"])), key=lambda i: c["tpd_ms_per_shape"][i])
        c["tpd_best_shape"] = shape_desc(best)
        c["tpd_best_ms"] = c["tpd_ms_per_shape"][best]
        c["rocprim_over_tpd"] = round(c["rocprim_ms"] / c["tpd_best_ms"], 3)
        log(f"[bench] sort vs rocPRIM: {c['pairs']} pairs, {c['key_bits']} bits: rocPRIM {c['rocprim_ms']} ms, "
            f"this sort {c['tpd_best_ms']} ms ({c['tpd_best_shape']})")
    return {"sort_vs_rocprim": cases}


def cpu_baseline(g, views, budget_s=12.0):
    """The oracle (multithreaded C restatement) on the first views of the same 32-view workload,
    until about budget_s seconds of CPU work; views per second."""
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.render_timed(g[:1000], oracle.make_camera([0, 0, -8], [0, 0, 0], [0, -1, 0], 64, 64), 3,
                        nthreads=threads)  # warm-up
    done, pairs, sizes = 0, 0, []
    t0 = time.perf_counter()
    for (_, v, pv, fo, w, h) in views:
        cam = oracle.Camera(v, pv, fo, w, h)
        P, _, _, _ = oracle.render_timed(g, cam, 3, nthreads=threads)
        done += 1
        pairs += P
        sizes.append(f"{w}x{h}")
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 4), "unit": "views/s", "cores": threads, "kind": "port",
            "sample": f"the first {done} of the 32 views ({', '.join(sizes)}), {pairs} pairs, rendered one "
                      f"after another by the oracle's C restatement with {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sort-shape", type=int, default=-1,
                    help="fixed radix-pass block shape instead of autotuning (for profiler runs)")
    ap.add_argument("--only", choices=["multiview", "plain", "sort-load", "fwd-bwd", "multiscene", "pipelined",
                                       "hbm-ceiling", "run-scatter", "sort-vs-rocprim"],
                    default=None,
                    help="run one leg only (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); views and scenes would shard "
              "across GPUs with no exchange, which is not implemented here.", file=sys.stderr)
        sys.exit(2)

    from torpedo_amd import scenes
    from torpedo_amd.engine import GaussianEngine

    W, H = 1920, 1080
    eye = [0.0, 0.0, -8.0]
    tune = args.sort_shape < 0
    eng = GaussianEngine(device=0)
    if not tune:
        eng.setSortShape(args.sort_shape)
    if args.only == "sort-load":
        g2 = scenes.cluster_scene(1_000_000, seed=2, log_scale_mean=-3.0)
        dt2, P2, st2 = time_plain(eng, g2, eye, W, H, args.steps, args.warmup, tune=tune)
        ach, t_launch = onesweep_roofline(P2, st2["sort_passes"], st2["radix_passes"])
        copy_gbs = eng.measureCopy(256 << 20, 3)  # a plain copy of 256 MiB for counter calibration
        print(json.dumps({"workload": "clusters-1M-1920x1080 log-scale mean -3", "pairs": P2, "stage_ms": st2,
                          "onesweep_avg_launch_us": round(t_launch * 1e6, 2), "onesweep_gbs": round(ach, 1),
                          "copy_256MiB_gbs": round(copy_gbs, 1)}), flush=True)
        eng.close()
        return
    if args.only == "plain":
        g = scenes.cluster_scene(1_000_000, seed=2)
        dt, P, st = time_plain(eng, g, eye, W, H, args.steps, args.warmup, tune=tune)
        print(json.dumps({"workload": "clusters-1M-1920x1080", "ms_per_frame": round(1000.0 * dt / args.steps, 4),
                          "pairs": P, "stage_ms": st}), flush=True)
        eng.close()
        return
    if args.only == "fwd-bwd":
        print(json.dumps(time_fwd_bwd(eng, args.steps, args.warmup)), flush=True)
        eng.close()
        return
    if args.only == "multiscene":
        print(json.dumps(time_multiscene(eng, args.steps, args.warmup)), flush=True)
        eng.close()
        return
    if args.only == "hbm-ceiling":
        print(json.dumps(hbm_ceiling(eng)), flush=True)
        eng.close()
        return
    if args.only == "run-scatter":
        print(json.dumps(run_scatter(eng)), flush=True)
        eng.close()
        return
    if args.only == "sort-vs-rocprim":
        eng.close()
        print(json.dumps(sort_vs_rocprim()), flush=True)
        return
    if args.only == "pipelined":
        g = scenes.cluster_scene(1_000_000, seed=2)
        print(json.dumps(time_pipelined(g, max(args.steps, 8), max(args.warmup, 2))), flush=True)
        eng.close()
        return

    log("[bench] generating clusters-1M scene")
    g = scenes.cluster_scene(1_000_000, seed=2)
    log("[bench] 32-view batched vs plain")
    mv = time_multiview(eng, g, args.steps, args.warmup, tune=tune)
    log(f"[bench] multiview: {mv}")
    if args.only == "multiview":
        mv["copy_256MiB_gbs"] = round(eng.measureCopy(256 << 20, 3), 1)  # counter calibration kernel
        print(json.dumps(mv), flush=True)
        eng.close()
        return

    log("[bench] plain 1080p frames")
    dt, P, st = time_plain(eng, g, eye, W, H, args.steps, args.warmup, tune=tune)
    ms_plain = 1000.0 * dt / args.steps
    log(f"[bench] plain: {ms_plain:.3f} ms/frame, P={P}, stages {st}")

    log("[bench] sort-load variant (log-scale mean -3)")
    g2 = scenes.cluster_scene(1_000_000, seed=2, log_scale_mean=-3.0)
    _, P2, st2 = time_plain(eng, g2, eye, W, H, max(5, args.steps // 2), 2, tune=tune)
    del g2
    log(f"[bench] sort-load: P={P2}, stages {st2}")

    # Roofline of the radix passes (k_onesweep) of the batched frame: the items are sorted by depth
    # (64-bit keys: the depth key and the carried emission record; 32-bit values: 12 bytes per
    # element read and 12 written per pass; the first pass reads every item's key, its value being
    # the item's index, and writes the visible items, the later passes read and write the visible
    # ones), then the pairs by tile (32-bit keys and values: 16 bytes); achieved = all passes' bytes
    # / all passes' time (the time-weighted mean over the frame's launches).
    es, so = mv["eager_stage_ms"], mv["sorts"]
    tuned = mv.get("autotune") or {}
    item_shape = tuned.get("item_shape", args.sort_shape)
    pair_shape = tuned.get("pair_shape", args.sort_shape)
    kept = so.get("items_kept", so["items"])
    item_bytes = (8.0 * so["items"] + 12.0 * kept + 24.0 * kept * (so["item_passes"] - 1)
                  if so["item_passes"] else 0.0)
    pair_bytes = 16.0 * so["pairs"] * so["pair_passes"]
    pass_bytes = item_bytes + pair_bytes
    t_passes = (so["item_passes_ms"] + so["pair_passes_ms"]) / 1000.0
    launches = so["item_passes"] + so["pair_passes"]
    ach = pass_bytes / t_passes / 1e9 if t_passes > 0 else 0.0
    traffic = ONESWEEP_TRAFFIC.get((item_shape, pair_shape))
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic["bytes_per_launch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "kernel": "k_onesweep (item sort: 64-bit keys; pair sort: 32-bit keys)",
                "sort_shape": {"items": shape_desc(item_shape), "pairs": shape_desc(pair_shape)},
                "items": so["items"], "items_kept": kept, "item_passes": so["item_passes"],
                "pairs": so["pairs"], "pair_passes": so["pair_passes"],
                "item_passes_gbs": round(item_bytes / (so["item_passes_ms"] / 1000.0) / 1e9, 1)
                if so["item_passes_ms"] > 0 else 0.0,
                "pair_passes_gbs": round(pair_bytes / (so["pair_passes_ms"] / 1000.0) / 1e9, 1)
                if so["pair_passes_ms"] > 0 else 0.0,
                "algorithmic_bytes_per_launch": round(pass_bytes / max(launches, 1)),
                "avg_launch_us": round(t_passes / max(launches, 1) * 1e6, 2),
                "workload": "multiview-32-clusters-1M (batched frame, issued eagerly for stage events)"}
    # The sort metric as the benchmark defines it: 24 B x P x ceil(K / d) over the time to order the
    # pairs by (tile, depth): both sorts, histograms included (K = global tile bits + 27 depth bits).
    t_sort = (es["item_sort"] + es["pair_sort"]) / 1000.0
    K = int(np.ceil(np.log2(max(sum(((v[0] + 15) // 16) * ((v[1] + 15) // 16) for v in mv["resolutions"]), 2)))) + 27
    d = SORT_SHAPES[pair_shape][0] if pair_shape is not None and 0 <= pair_shape < len(SORT_SHAPES) else 9
    sort_metric = 24.0 * mv["pairs"] * ((K + d - 1) // d) / t_sort / 1e9 if t_sort > 0 else 0.0
    roofline["sort_ms"] = round(t_sort * 1000.0, 4)
    roofline["sort_gkeys_per_s"] = round(mv["pairs"] / t_sort / 1e9, 3) if t_sort > 0 else 0.0
    ach2, t2 = onesweep_roofline(P2, st2["sort_passes"], st2["radix_passes"])
    ach1, _ = onesweep_roofline(P, st["sort_passes"], st["radix_passes"])

    log("[bench] forward + backward, 3M Gaussians 1920x1080")
    fb = time_fwd_bwd(eng, max(5, args.steps // 2), 2)
    log(f"[bench] fwd+bwd: {fb}")
    log("[bench] 16 scenes x 4 views training step")
    ms5 = time_multiscene(eng, max(5, args.steps // 2), 2)
    log(f"[bench] multiscene: {ms5}")
    log("[bench] frame pipelining, 1920x1080 single-view frames, two contexts")
    pipe = time_pipelined(g, max(16, args.steps), 4)
    log(f"[bench] pipelined: {pipe}")
    copy_gbs = eng.measureCopy(1 << 30, 10)
    log(f"[bench] plain copy: {copy_gbs:.0f} GB/s")
    # Blend: ~22 FP32 operations + 1 exp per evaluated pixel-splat pair (device-counted), on the
    # batched frame (k_batch_blend) and on the plain 1080p frame (k_blend).
    def blend_roofline(ev, ms, kernel, workload):
        tflops = 22.0 * ev / (ms / 1000.0) / 1e12 if ms > 0 else 0.0
        return {"bound": "valu", "achieved": round(tflops, 3), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / VALU_PEAK_TFLOPS, 4), "evaluations": ev, "blend_ms": round(ms, 4),
                "kernel": kernel, "workload": workload,
                "note": "22 flop per evaluated pixel-splat pair (plus one exp), vector FP32 peak"}
    roofline_blend = blend_roofline(mv["blend_evaluations"], es["blend"], "k_batch_blend",
                                    "multiview-32-clusters-1M batched frame")
    roofline_blend["plain_1080p"] = blend_roofline(st["blend_evaluations"], st["blend"], "k_blend",
                                                   "clusters-1M-1920x1080 plain frame")

    out = {
        "metric": METRIC,
        "value": mv["batched_views_per_s"],
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mv["batched_ms_per_frame_wall"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scenes and camera rigs, no dataset)",
        "config": {"workload": "multiview-32-clusters-1M", "gaussians": 1_000_000, "views": 32,
                   "radius": 8.0, "sh_degree": 3, "tile": 16, "pairs": mv["pairs"],
                   "sort_shape_batch": {"items": shape_desc((mv.get("autotune") or {}).get("item_shape", args.sort_shape)),
                                        "pairs": shape_desc((mv.get("autotune") or {}).get("pair_shape", args.sort_shape))},
                   "sort_shape_plain": shape_desc((mv.get("autotune_plain") or {}).get("best_shape", args.sort_shape)),
                   "parallelism": "single GPU", "graph": "one recorded HIP graph per frame"},
        "multiview_32": mv,
        "plain_1080p": {"ms_per_frame": round(ms_plain, 4), "views_per_s": round(1000.0 / ms_plain, 2), "pairs": P,
                        "stage_ms": {k: round(v, 4) for k, v in st.items()
                                     if k not in ("radix_passes", "autotune", "blend_evaluations")},
                        "autotune": st.get("autotune")},
        "sort_hbm_frac": {"multiview_32": round(sort_metric / HBM_PEAK_GBS, 4), "pairs_multiview_32": mv["pairs"],
                          "multiview_32_key_bits": K,
                          "sort_load_variant": round(ach2 / HBM_PEAK_GBS, 4), "pairs_sort_load": P2,
                          "plain_1080p": round(ach1 / HBM_PEAK_GBS, 4), "pairs_plain_1080p": P,
                          "gkeys_per_s": {"multiview_32": roofline["sort_gkeys_per_s"],
                                          "sort_load_variant": round(sort_keys_per_s(P2, st2) / 1e9, 3),
                                          "plain_1080p": round(sort_keys_per_s(P, st) / 1e9, 3)}},
        "roofline": roofline,
        "roofline_blend": roofline_blend,
        "fwd_bwd_3M_1080p": fb,
        "multiscene_16x4_step": ms5,
        "pipelined_1080p": pipe,
        "copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle restatement)")
        out["cpu_baseline"] = cpu_baseline(g, multiview_views())
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
