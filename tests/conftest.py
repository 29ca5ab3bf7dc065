import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libtgvx.so")


def _gpu_available():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def ctx():
    if not _gpu_available():
        pytest.fail("no GPU visible: gpu-marked tests must run on the GPU box")
    from tgvreconclaude_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def sphere():
    """The CPU-sized workload: unit sphere, 24 ring views 320x240, depth <= 7, 10 bins."""
    from tgvreconclaude_amd import scene
    import oracle as O
    sc = scene.sphere_scene()
    vs, keep = O.views_from_scene(sc)
    hist = O.histograms(sc.cubes, vs, sc.origin, sc.r_root, 10)
    return sc, hist
