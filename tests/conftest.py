import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# torch ships its own HIP runtime under an unversioned file name; loading it before libtpd_hip.so
# lets the library bind to that same runtime (soname libamdhip64.so.7) instead of a second copy.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def c1_camera(orc, w=1280, h=720, dist=3.0):
    return orc.make_camera([0, 0, -dist], [0, 0, 0], [0, -1, 0], w, h)


def small_scene(n, seed, spread=1.0, log_scale_mean=-4.5):
    from torpedo_amd import scenes
    g = scenes.uniform_scene(n, seed=seed, log_scale_mean=log_scale_mean)
    g[:, 0:3] *= spread
    return g


def report_fp64(name, img32, img64):
    """States one image's error against the FP64 twin (printed; also appended as a JSON line to the
    file named by TPD_FP64_LOG when that is set) and returns (max, mean, 99.9th percentile)."""
    import json
    err = np.abs(np.asarray(img32, np.float64) - np.asarray(img64, np.float64))
    stats = (float(err.max()), float(err.mean()), float(np.quantile(err, 0.999)))
    print(f"[fp64] {name}: image vs FP64 twin max-abs {stats[0]:.3g}, mean {stats[1]:.3g}, "
          f"p99.9 {stats[2]:.3g}", flush=True)
    path = os.environ.get("TPD_FP64_LOG")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"image": name, "shape": list(np.shape(img32)), "max_abs": stats[0],
                                "mean_abs": stats[1], "p999_abs": stats[2]}) + "\n")
    return stats
