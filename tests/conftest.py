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
