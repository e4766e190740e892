"""The synthetic depth generator (stands in for the reference's frame loader): seeded
determinism against committed digests, analytic depth, the noise and dropout model."""
import hashlib
import json
import os

import numpy as np

from conftest import ROOT


def test_golden_digests(vh, wl):
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "synth_room_small.json")))
    scene = vh.Scene(g["preset"], g["seed"])
    for f, want in g["frames"].items():
        d = scene.depth(int(f))
        assert hashlib.sha256(d.tobytes()).hexdigest() == want["depth_sha256"], f
        assert hashlib.sha256(scene.pose(int(f)).tobytes()).hexdigest() == want["pose_sha256"], f


def test_pose_path_and_frame_independence(vh, wl):
    scene = vh.Scene(1, wl.SEED)
    a, b = scene.pose(10), scene.pose(11)
    step = np.linalg.norm(a[:3, 3] - b[:3, 3])
    assert abs(step - 0.02) < 1e-4  # 2 cm per frame
    cosang = (np.trace(a[:3, :3].T.astype(np.float64) @ b[:3, :3]) - 1) / 2
    assert abs(np.degrees(np.arccos(np.clip(cosang, -1, 1))) - 1.0) < 1e-3  # 1 degree per frame
    R = a[:3, :3].astype(np.float64)
    assert np.abs(R.T @ R - np.eye(3)).max() < 1e-6 and abs(np.linalg.det(R) - 1) < 1e-6
    d_range, _ = scene.frames(5, 3)
    assert np.array_equal(d_range[1], scene.depth(6))  # any frame reproduces independently


def test_clean_depth_of_a_wall(vh, wl):
    """Camera 1 m in front of the x = 0 wall, looking at it: depth is exactly 1 m at the
    centre pixel, and 1 m everywhere the wall fills the view (fronto-parallel plane)."""
    scene = vh.Scene(1, wl.SEED)
    T = np.eye(4, dtype=np.float32)
    # camera axes: x right = world -y, y down = world -z, z forward = world -x
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = (0, -1, 0), (0, 0, -1), (-1, 0, 0), (1.0, 1.5, 1.25)
    d = scene.clean_depth(0, T)
    assert abs(d[240, 320] - 1.0) < 1e-6
    centre = d[200:280, 280:360]
    # primitives may stand in front of the wall; where the wall is seen, depth is 1 m
    assert np.isclose(centre, 1.0, atol=1e-6).mean() > 0.5


def test_noise_and_dropout_model(vh, wl):
    scene = vh.Scene(2, wl.SEED)
    T = scene.pose(3)
    clean = scene.clean_depth(3, T).astype(np.float64)
    noisy = scene.depth(3, T).astype(np.float64) * 1e-3
    valid = clean > 0
    drop = valid & (noisy == 0)
    assert 0.008 < drop.sum() / valid.sum() < 0.012  # 1% dropout
    m = valid & (noisy > 0)
    z = clean[m]
    sigma = 0.0012 + 0.0019 * (z - 0.4) ** 2
    e = (noisy[m] - z) / np.sqrt(sigma ** 2 + 0.0005 ** 2 / 3)  # + mm quantisation
    assert abs(e.mean()) < 0.02 and abs(e.std() - 1.0) < 0.05
