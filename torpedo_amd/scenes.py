"""Deterministic synthetic Gaussian scenes and camera rigs for the benchmark workloads.

Every random draw comes from SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
number generators", OOPSLA 2014) evaluated counter-style with numpy, one independent stream per
(seed, attribute); normals use the Box-Muller transform and unit quaternions Shoemake's method
("Uniform random rotations", Graphics Gems III, 1992).  A scene is an (N, 60) float32 array in
the 240-byte GaussianPoint layout of torpedo/volumetric/include/torpedo/volumetric/
GaussianGeometry.h:10-30: position xyz, opacity (activated), quaternion xyzw, scale xyz
(activated) + modifier w = 1, sh[48] (DC rgb, then 15 coefficients per channel).

Workloads (sizes from the benchmark definitions):
  uniform   N Gaussians, means U[-1,1]^3                      (100k @ 1280x720, eye distance 3)
  clusters  64 anisotropic clusters, centres U[-5,5]^3,       (1M / 3M @ 1920x1080, eye distance 8)
            per-cluster per-axis std U(0.1,1), random rotation
Attributes: log-scale ~ N(mu, 0.5) per axis (mu = -4.5, or -3.0 for the sort-load variant),
opacity = sigmoid(N(0,1)), sh DC ~ N(0.5,0.2), other sh ~ N(0,0.2).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_CHUNK = 1 << 22
_THREADS = max(1, min(16, os.cpu_count() or 1))


def _parallel(fn, n: int, chunk: int) -> None:
    """Runs fn(lo, hi) over [0, n) in chunks on a thread pool (numpy releases the GIL); the
    chunking changes no value."""
    spans = [(lo, min(n, lo + chunk)) for lo in range(0, n, chunk)]
    if len(spans) <= 1 or _THREADS == 1:
        for lo, hi in spans:
            fn(lo, hi)
        return
    with ThreadPoolExecutor(_THREADS) as ex:
        list(ex.map(lambda sp: fn(*sp), spans))


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_base(seed: int, attr: int) -> np.uint64:
    with np.errstate(over="ignore"):
        s = _mix(np.array([np.uint64(seed) * GOLDEN + np.uint64(attr) * np.uint64(0xD1B54A32D192ED03)],
                          dtype=np.uint64))[0]
    return s


def uniforms(seed: int, attr: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from SplitMix64 stream (seed, attr): u_i = mix(base + (i+1)*golden) >> 11 * 2^-53."""
    base = _stream_base(seed, attr)
    out = np.empty(n, np.float64)

    def run(lo, hi):
        with np.errstate(over="ignore"):
            i = np.arange(lo + 1, hi + 1, dtype=np.uint64)
            z = _mix(base + i * GOLDEN)
            out[lo:hi] = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    _parallel(run, n, _CHUNK)
    return out


def normals(seed: int, attr: int, n: int) -> np.ndarray:
    """n standard normals by Box-Muller over pairs of uniforms of stream (seed, attr)."""
    m = (n + 1) // 2
    u = uniforms(seed, attr, 2 * m)
    out = np.empty(2 * m, np.float64)

    def run(lo, hi):  # pairs [lo, hi)
        u1 = 1.0 - u[2 * lo:2 * hi:2]
        u2 = u[2 * lo + 1:2 * hi:2]
        r = np.sqrt(-2.0 * np.log(u1))
        out[2 * lo:2 * hi:2] = r * np.cos(2.0 * math.pi * u2)
        out[2 * lo + 1:2 * hi:2] = r * np.sin(2.0 * math.pi * u2)

    _parallel(run, m, _CHUNK // 2)
    return out[:n]


def unit_quaternions(seed: int, attr: int, n: int) -> np.ndarray:
    """Shoemake: (x, y, z, w) uniform on S^3."""
    u = uniforms(seed, attr, 3 * n).reshape(n, 3)
    a = np.sqrt(1.0 - u[:, 0])
    b = np.sqrt(u[:, 0])
    t1 = 2.0 * math.pi * u[:, 1]
    t2 = 2.0 * math.pi * u[:, 2]
    return np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], axis=1)


def _quat_to_rot(q: np.ndarray) -> np.ndarray:
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def attributes(seed: int, n: int, log_scale_mean: float = -4.5, out: np.ndarray | None = None) -> np.ndarray:
    """Fills columns 3..59 (everything but the means) of an (n, 60) float32 record array."""
    g = np.zeros((n, 60), np.float32) if out is None else out
    g[:, 3] = (1.0 / (1.0 + np.exp(-normals(seed, 2, n)))).astype(np.float32)
    g[:, 4:8] = unit_quaternions(seed, 3, n).astype(np.float32)
    g[:, 8:11] = np.exp(log_scale_mean + 0.5 * normals(seed, 4, 3 * n)).reshape(n, 3).astype(np.float32)
    g[:, 11] = 1.0
    g[:, 12:15] = (0.5 + 0.2 * normals(seed, 5, 3 * n)).reshape(n, 3).astype(np.float32)

    def sh_rest(lo, hi):
        g[lo:hi, 15:60] = (0.2 * normals(seed * 1000003 + lo, 6, 45 * (hi - lo))).reshape(hi - lo, 45).astype(
            np.float32)

    _parallel(sh_rest, n, _CHUNK // 16)
    return g


def uniform_scene(n: int, seed: int = 1, log_scale_mean: float = -4.5) -> np.ndarray:
    """Means uniform in [-1, 1]^3."""
    g = np.zeros((n, 60), np.float32)
    g[:, 0:3] = (2.0 * uniforms(seed, 1, 3 * n) - 1.0).reshape(n, 3).astype(np.float32)
    return attributes(seed, n, log_scale_mean, g)


def cluster_scene(n: int, seed: int = 2, clusters: int = 64, log_scale_mean: float = -4.5) -> np.ndarray:
    """Mixture of anisotropic Gaussian clusters: centres U[-5,5]^3, per-axis std U(0.1,1), random rotation,
    uniform membership."""
    g = np.zeros((n, 60), np.float32)
    cu = uniforms(seed, 10, clusters * 6).reshape(clusters, 6)
    centres = -5.0 + 10.0 * cu[:, 0:3]
    stds = 0.1 + 0.9 * cu[:, 3:6]
    rots = _quat_to_rot(unit_quaternions(seed, 11, clusters))
    member = np.minimum((uniforms(seed, 12, n) * clusters).astype(np.int64), clusters - 1)
    local = normals(seed, 1, 3 * n).reshape(n, 3) * stds[member]
    g[:, 0:3] = (np.einsum("nij,nj->ni", rots[member], local) + centres[member]).astype(np.float32)
    return attributes(seed, n, log_scale_mean, g)


def fibonacci_sphere(count: int, radius: float, seed: int = 4):
    """count camera positions on a Fibonacci sphere (poles excluded), looking at the origin, up = +z."""
    i = np.arange(count, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * i / count
    z = np.clip(z, -0.95, 0.95)
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([radius * r * np.cos(phi), radius * r * np.sin(phi), radius * z], 1).astype(np.float32)


RESOLUTIONS = [(256, 256), (512, 512), (640, 480), (800, 600), (1024, 768), (1280, 720), (1600, 900), (1920, 1080)]


def multiview_rig(count: int = 32, radius: float = 8.0, seed: int = 4):
    """(eye, (w, h)) for the batched multi-view workload: resolutions drawn from RESOLUTIONS."""
    eyes = fibonacci_sphere(count, radius, seed)
    pick = np.minimum((uniforms(seed, 20, count) * len(RESOLUTIONS)).astype(np.int64), len(RESOLUTIONS) - 1)
    return [(eyes[k], RESOLUTIONS[pick[k]]) for k in range(count)]


def multiscene_views(count: int = 16, per_scene: int = 4, radius: float = 8.0, seed: int = 5):
    """Per scene s: per_scene (eye, (w, h)) with eyes on a Fibonacci sphere (seed + s) and
    W ~ U{512..1600}, H ~ U{512..900}."""
    out = []
    for s in range(count):
        eyes = fibonacci_sphere(per_scene, radius, seed + s)
        uw, uh = uniforms(seed + s, 40, per_scene), uniforms(seed + s, 41, per_scene)
        out.append([(eyes[k], (512 + min(int(uw[k] * 1089), 1088), 512 + min(int(uh[k] * 389), 388)))
                    for k in range(per_scene)])
    return out


def multiscene_sizes(count: int = 16, seed: int = 5):
    """Gaussian counts log-uniform in [5e4, 2e6]."""
    u = uniforms(seed, 30, count)
    return [int(math.floor(math.exp(math.log(5e4) + u[k] * (math.log(2e6) - math.log(5e4))))) for k in range(count)]
