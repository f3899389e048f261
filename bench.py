"""Benchmark of the Gaussian rasterization path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

A step is one plain forward frame (GaussianEngine::rasterFrame equivalent: preprocess, scan +
duplicate, radix sort, ranges, blend, synchronised) of the workload "clusters-1M-1920x1080":
1M Gaussians in 64 anisotropic clusters (seeded, synthetic), camera at distance 8, SH degree 3,
16x16 tiles.  `value` is frames (views) per second.  The sort-load variant (log-scale mean -3,
~10-20M pairs) is run for the sort's HBM fraction.  Stage times come from HIP events recorded on
the render stream inside the timed frames.

One JSON line is printed.  Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "views/s batched vs plain; fwd+bwd ms @3M Gaussians 1080p; sort % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def camera_uniforms(eye, w, h):
    from torpedo_amd.engine import PerspectiveCamera
    cam = PerspectiveCamera(w, h)
    cam.lookAt(eye, [0, 0, 0], [0, -1, 0])
    view, pv, fo = cam.uniforms()
    return view, pv, fo


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
    med = {k: float(np.median([s[k] for s in stages])) for k in stages[0]}
    if tuned is not None:
        med["autotune"] = tuned
    return dt, P, med


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
    med = {k: round(float(np.median([r[k] for r in rec])), 4) for k in rec[0]}
    return {"ms": med["fwd_bwd_span"], "wall_ms_per_step": round(wall, 4), "pairs": P, "gaussians": n,
            "width": W, "height": H, "stage_ms": {k: v for k, v in med.items() if k not in ("radix_passes",)}}


def cpu_baseline(g, eye, w, h):
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    cam = oracle.make_camera(eye, [0, 0, 0], [0, -1, 0], w, h)
    oracle.render_timed(g[:1000], cam, 3, nthreads=threads)  # warm-up
    t0 = time.perf_counter()
    P, _, _, st = oracle.render_timed(g, cam, 3, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "views/s", "cores": threads, "kind": "port",
            "sample": f"1 frame of the same workload (P={P}); stage s: "
                      + ", ".join(f"{k}={v:.3f}" for k, v in zip(
                          ["preprocess", "scan+dup", "sort", "ranges", "blend"], st))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sort-shape", type=int, default=-1,
                    help="fixed radix-pass block shape instead of autotuning (for profiler runs)")
    ap.add_argument("--only-sort-load", action="store_true",
                    help="run only the sort-load frames (for profiler runs of the radix pass)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); views and scenes would shard "
              "across GPUs with no exchange, which this build does not run.", file=sys.stderr)
        sys.exit(2)

    from torpedo_amd import scenes
    from torpedo_amd.engine import GaussianEngine

    W, H = 1920, 1080
    eye = [0.0, 0.0, -8.0]
    tune = args.sort_shape < 0
    eng = GaussianEngine(device=0)
    if not tune:
        eng.setSortShape(args.sort_shape)
    if args.only_sort_load:
        g2 = scenes.cluster_scene(1_000_000, seed=2, log_scale_mean=-3.0)
        dt2, P2, st2 = time_plain(eng, g2, eye, W, H, args.steps, args.warmup, tune=tune)
        t_launch = st2["sort_passes"] / 1000.0 / st2["radix_passes"]
        # A plain copy of 256 MiB (read + write = 512 MiB per launch) for counter calibration.
        copy_gbs = eng.measureCopy(256 << 20, 3)
        print(json.dumps({"workload": "clusters-1M-1920x1080 log-scale mean -3", "pairs": P2, "stage_ms": st2,
                          "onesweep_avg_launch_us": round(t_launch * 1e6, 2),
                          "onesweep_gbs": round(24.0 * P2 / t_launch / 1e9, 1),
                          "copy_256MiB_gbs": round(copy_gbs, 1)}), flush=True)
        eng.close()
        return
    log("[bench] generating clusters-1M scene")
    g = scenes.cluster_scene(1_000_000, seed=2)
    log("[bench] timing plain forward frames")
    dt, P, st = time_plain(eng, g, eye, W, H, args.steps, args.warmup, tune=tune)
    ms = 1000.0 * dt / args.steps
    log(f"[bench] plain: {ms:.3f} ms/frame, P={P}, stages {st}")

    log("[bench] sort-load variant (log-scale mean -3)")
    g2 = scenes.cluster_scene(1_000_000, seed=2, log_scale_mean=-3.0)
    dt2, P2, st2 = time_plain(eng, g2, eye, W, H, max(5, args.steps // 2), 2, tune=tune)
    log(f"[bench] sort-load: P={P2}, stages {st2}")

    # Roofline of the radix pass (onesweep) kernel: 24 algorithmic bytes per pair per pass
    # (key + value read and written), average launch = all passes / pass count.
    passes = st2["radix_passes"]
    t_launch_s = st2["sort_passes"] / 1000.0 / passes
    achieved = 24.0 * P2 / t_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_onesweep", "pairs": P2, "passes": passes,
                "avg_launch_us": round(t_launch_s * 1e6, 2)}
    t_launch_c2 = st["sort_passes"] / 1000.0 / max(st["radix_passes"], 1)
    sort_frac_c2 = 24.0 * P / t_launch_c2 / 1e9 / HBM_PEAK_GBS if P else 0.0
    log("[bench] forward + backward, 3M Gaussians 1920x1080")
    fb = time_fwd_bwd(eng, max(5, args.steps // 2), 2)
    log(f"[bench] fwd+bwd: {fb}")
    copy_gbs = eng.measureCopy(1 << 30, 10)
    log(f"[bench] plain copy: {copy_gbs:.0f} GB/s")
    # Blend: ~22 FP32 operations + 1 exp per evaluated pixel-splat pair (device-counted).
    ev = st["blend_evaluations"]
    blend_tflops = 22.0 * ev / (st["blend"] / 1000.0) / 1e12
    roofline_blend = {"bound": "valu", "achieved": round(blend_tflops, 3), "peak": 157.3, "unit": "TFLOP/s",
                      "frac": round(blend_tflops / 157.3, 4), "evaluations": ev, "kernel": "k_blend",
                      "note": "22 flop per evaluated pixel-splat pair (plus one exp), vector FP32 peak"}

    out = {
        "metric": METRIC,
        "value": round(args.steps / dt, 3),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene, no dataset)",
        "config": {"workload": "clusters-1M-1920x1080 plain forward", "gaussians": 1_000_000, "width": W,
                   "height": H, "sh_degree": 3, "tile": 16, "pairs": P, "sort_digit_bits": 8,
                   "parallelism": "single GPU"},
        "stage_ms": {k: round(v, 4) for k, v in st.items()
                     if k not in ("radix_passes", "autotune", "blend_evaluations")},
        "autotune": {"clusters_1M": st.get("autotune"), "sort_load_variant": st2.get("autotune")},
        "sort_hbm_frac": {"sort_load_variant": round(achieved / HBM_PEAK_GBS, 4), "pairs": P2,
                          "clusters_1M": round(sort_frac_c2, 4), "pairs_clusters_1M": P},
        "roofline": roofline,
        "roofline_blend": roofline_blend,
        "fwd_bwd_3M_1080p": fb,
        "copy_gbs": round(copy_gbs, 1),
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (oracle restatement, one frame)")
        out["cpu_baseline"] = cpu_baseline(g, eye, W, H)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
