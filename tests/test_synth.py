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


def test_rgb_frames_are_a_world_checkerboard_of_surface_colours(vh):
    """vh_scene_render_rgb: each surface has one base colour (channels 48..239), dark cells
    carry floor(5/8) of it, and the cell parity of a pixel's colour is the parity of its
    world point (from the clean depth and the pose) on the 10 cm lattice."""
    sc = vh.Scene(1, 20240611)
    f = 33  # three surfaces in view
    T = sc.pose(f).astype(np.float64)
    rgb = sc.rgb(f).astype(np.int64)
    z = sc.clean_depth(f).astype(np.float64)
    H, W = z.shape
    fx, fy, cx, cy = sc.intr
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    d = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(z)], -1)
    p = d @ T[:3, :3].T * z[..., None] + T[:3, 3]
    # coordinates on a room face lie on a cell plane: they count as that plane's cell; the
    # others must be clear of cell planes (the FP32 depth moves a point by ~1e-6 m)
    q = p * 10.0
    ext = np.array(sc.extent, np.float64) * 10.0
    on_face = (np.abs(q) < 1e-3) | (np.abs(q - ext) < 1e-3)
    dist = np.abs(q - np.rint(q))
    clear = (on_face | (dist > 1e-3)).all(-1) & (on_face.sum(-1) <= 1) & (z > 0)  # not at room edges
    cell = np.where(on_face, np.rint(q), np.floor(q)).astype(np.int64).sum(-1)
    key = rgb[..., 0] << 16 | rgb[..., 1] << 8 | rgb[..., 2]
    colours = np.unique(key[clear])
    assert 2 <= len(colours) <= 2 * (6 + sc.n_spheres + sc.n_boxes)

    def darken(c):
        return ((c >> 16) * 5 // 8) << 16 | (((c >> 8) & 255) * 5 // 8) << 8 | (c & 255) * 5 // 8

    # base colours are the ones whose dark shade also appears (both cell parities are seen)
    seen = {int(c) for c in colours}
    light = {c for c in seen if darken(c) in seen and all(48 <= (c >> s) & 255 <= 239 for s in (0, 8, 16))}
    dark = {darken(c) for c in light}
    assert len(light) >= 3 and not (light & dark) and seen == light | dark
    is_dark = np.isin(key, list(dark))
    assert np.array_equal(is_dark[clear], (cell[clear] & 1) == 1)
    # a pose moves the image but not the colours of the world
    assert np.array_equal(sc.rgb(f, sc.pose(f)), sc.rgb(f))
