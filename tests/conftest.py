import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X; the HIP path of libvh is exercised")


@pytest.fixture(scope="session")
def vh():
    mod, _ = __graft_entry__.package()
    mod.load()
    return mod


@pytest.fixture(scope="session")
def wl():
    return __graft_entry__.package()[1]


@pytest.fixture(scope="session")
def oracle():
    import oracle as o

    o.load()
    return o


@pytest.fixture(scope="session")
def small_frames(vh, wl):
    """The 40 frames of the small-room workload (uint16 depth, poses)."""
    w = wl.WORKLOADS["room_small_1cm"]
    scene = vh.Scene(w["scene"], wl.SEED)
    return scene.frames(0, w["frames"])


def plane_depth(w, D_mm, H=None, W=None):
    H = H or w["height"]
    W = W or w["width"]
    return np.full((H, W), D_mm, np.uint16)


def look_down_z(tx=0.0, ty=0.0, tz=0.0):
    """Camera at (tx,ty,tz) looking along +z of the world (identity rotation)."""
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = (tx, ty, tz)
    return T
