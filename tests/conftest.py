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


def sphere_views(w, radius=0.25, dist=0.9, scale=1e-4):
    """Six views of a sphere of `radius` (metres) at the origin, from distance `dist` along
    +-x, +-y, +-z, rendered analytically (ray-sphere intersection in FP64) into uint16 depth
    of `scale` metres per unit (0 = background). Returns (depth (6,H,W) uint16, c2w (6,4,4))."""
    H, W = w["height"], w["width"]
    fx, fy, cx, cy = w["fx"], w["fy"], w["cx"], w["cy"]
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    dirs_cam = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], -1)
    depths, poses = [], []
    for axis in range(3):
        for sign in (1.0, -1.0):
            P = np.zeros(3)
            P[axis] = sign * dist
            f = -P / np.linalg.norm(P)
            up = np.array([0.0, 0.0, 1.0]) if axis != 2 else np.array([0.0, 1.0, 0.0])
            r = np.cross(f, up)
            r /= np.linalg.norm(r)
            d = np.cross(f, r)
            R = np.stack([r, d, f], 1)  # columns: camera x (right), y (down), z (forward)
            dw = dirs_cam @ R.T
            b = dw @ P
            a = (dw * dw).sum(-1)
            c = P @ P - radius * radius
            disc = b * b - a * c
            t = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / a, 0.0)
            depths.append(np.where(t > 0, np.rint(t / scale), 0).astype(np.uint16))
            T = np.eye(4, dtype=np.float32)
            T[:3, :3] = R
            T[:3, 3] = P
            poses.append(T)
    return np.stack(depths), np.stack(poses)
