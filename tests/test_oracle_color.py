"""Pins of the oracle's colour channel (CPU only).

The reference keeps an 8-bit colour per voxel, averaged with the same weights as the TSDF
(src/tsdf.cu:743-745: (c * w_old + rgb) / w_new into an unsigned char), and interpolates
vertex colours along the cube edge like positions (src/tsdf.cu:1655-1657). The checks
below fix the oracle's restatement by facts independent of its code: the FP32 quotient of
two integers below 2^24 truncated to 8 bits is the integer floor quotient (the gap between
a non-integer quotient and the next integer, >= 1/(w+1) >= 2^-16, exceeds half an FP32
ulp of any value below 256); a constant colour is a fixed point of the average; a single
frame gives every observed voxel the colour of the pixel its centre projects to (brute
force in FP64); a vertex colour lies on the segment between its corners' colours at the
vertex's own position along the edge.
"""
import numpy as np
import pytest

from conftest import look_down_z, plane_depth


def test_color_update_is_the_integer_floor_mean(oracle):
    rng = np.random.default_rng(5)
    cases = [(0, 0.0, 0), (255, 0.0, 0), (0, 0.0, 255), (255, 65535.0, 255), (0, 65535.0, 255),
             (254, 1.0, 255), (1, 2.0, 0), (128, 3.0, 127)]
    cases += [(int(c), float(w), int(p)) for c, w, p in
              zip(rng.integers(0, 256, 3000), rng.integers(0, 65536, 3000), rng.integers(0, 256, 3000))]
    for c, w, p in cases:
        px = np.array([p, 255 - p, (p * 7) & 255], np.uint8)
        got = oracle.color_update([c, c, 255 - c], w, px)
        want = [(c * int(w) + int(px[0])) // (int(w) + 1), (c * int(w) + int(px[1])) // (int(w) + 1),
                ((255 - c) * int(w) + int(px[2])) // (int(w) + 1)]
        assert list(got) == want, (c, w, p, list(got), want)


def test_first_observation_takes_the_pixel_and_constants_are_fixed_points(oracle):
    for p in (0, 1, 77, 254, 255):
        px = np.array([p, 255 - p, 9], np.uint8)
        assert list(oracle.color_update([200, 13, 4], 0.0, px)) == [p, 255 - p, 9]  # w_old = 0
        for w in (1.0, 2.0, 7.0, 254.0, 1000.0):
            assert list(oracle.color_update(px, w, px)) == list(px)


def test_vertex_colour_snaps_and_edge_position(oracle):
    a = np.array([0, 0, 0], np.float32)
    b = np.array([1, 0, 0], np.float32)
    ca = np.array([10, 200, 0], np.uint8)
    cb = np.array([250, 20, 255], np.uint8)
    # snaps: |va| < 1e-5 -> a, |vb| < 1e-5 -> b, |va - vb| < 1e-5 -> a
    p, c = oracle.vertex_interp_rgb(a, b, 5e-6, 0.5, ca, cb)
    assert np.array_equal(p, a) and np.array_equal(c, ca)
    p, c = oracle.vertex_interp_rgb(a, b, -0.5, -9e-6, ca, cb)
    assert np.array_equal(p, b) and np.array_equal(c, cb)
    p, c = oracle.vertex_interp_rgb(a, b, 0.3, 0.3 + 5e-6, ca, cb)
    assert np.array_equal(p, a) and np.array_equal(c, ca)
    # equal corner colours pass through unchanged
    p, c = oracle.vertex_interp_rgb(a, b, -0.25, 0.75, ca, ca)
    assert np.array_equal(c, ca)
    # a vertex at fraction x of the edge a->b (x = its coordinate here) has colour within one
    # truncation of ca + x (cb - ca), and always between the two corner colours
    rng = np.random.default_rng(3)
    for _ in range(2000):
        va = -rng.uniform(0.01, 1.0)
        vb = rng.uniform(0.01, 1.0)
        if rng.random() < 0.5:
            va, vb = -va, -vb
        c1 = rng.integers(0, 256, 3).astype(np.uint8)
        c2 = rng.integers(0, 256, 3).astype(np.uint8)
        p, c = oracle.vertex_interp_rgb(a, b, va, vb, c1, c2)
        x = float(p[0])
        assert 0.0 <= x <= 1.0
        exact = c1.astype(np.float64) + x * (c2.astype(np.float64) - c1.astype(np.float64))
        assert np.all(np.abs(c.astype(np.float64) - np.floor(exact)) <= 1.0), (va, vb, c1, c2, c, exact)
        lo = np.minimum(c1, c2)
        hi = np.maximum(c1, c2)
        assert np.all(c >= lo) and np.all(c <= hi)


def _texture(H, W, f=0):
    u = np.arange(W)[None, :]
    v = np.arange(H)[:, None]
    r = (u * 7 + f * 13) & 255
    g = (v * 5 + 3 * f) & 255
    b = ((u ^ v) + 11 * f) & 255
    return np.stack(np.broadcast_arrays(r, g, b), axis=-1).astype(np.uint8)


def test_single_frame_colours_are_the_projected_pixels(oracle, wl):
    """One frame of a plane at 1.2 m with a textured image: every observed voxel holds the
    colour of the pixel its centre projects to (FP64 projection, rounded; voxels whose
    projection lies within 1e-3 px of a rounding boundary are skipped)."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    H, W = w["height"], w["width"]
    T = look_down_z(0.013, -0.021, 0.0)
    depth = plane_depth(w, 1200)
    rgb = _texture(H, W)
    ref = oracle.Oracle(w, storage=0, weight_cap=0, color=True)
    ref.integrate(depth, T, rgb=rgb)
    keys, s, wt, _, _ = ref.export()
    k2, col = ref.export_rgb()
    assert np.array_equal(keys, k2)
    i = np.arange(512)
    lat = keys[:, None, :] * 8 + np.stack([i >> 6, (i >> 3) & 7, i & 7], -1)[None]
    pos = lat.astype(np.float64) * w["vox_size"] - T[:3, 3].astype(np.float64)
    m = wt > 0
    x, y, z = pos[..., 0][m], pos[..., 1][m], pos[..., 2][m]
    uf = w["fx"] * x / z + w["cx"]
    vf = w["fy"] * y / z + w["cy"]
    clear = (np.abs(uf - np.floor(uf) - 0.5) > 1e-3) & (np.abs(vf - np.floor(vf) - 0.5) > 1e-3)
    u = np.rint(uf).astype(np.int64)[clear]
    v = np.rint(vf).astype(np.int64)[clear]
    got = col[m][clear]
    assert clear.sum() > 100000
    assert np.array_equal(got, rgb[v, u]), f"{(got != rgb[v, u]).any(-1).sum()} voxels differ"
    assert np.all(col[~m] == 0)  # never observed


def test_two_frames_average_and_mesh_colours(oracle, wl):
    """Two frames of the same plane in constant colours p then q: observed voxels hold
    floor((p + q) / 2); every marching-cubes vertex of the plane takes that colour."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    H, W = w["height"], w["width"]
    T = look_down_z()
    depth = plane_depth(w, 1000)
    p = np.array([200, 10, 77], np.uint8)
    q = np.array([31, 255, 78], np.uint8)
    ref = oracle.Oracle(w, storage=0, weight_cap=0, color=True)
    ref.integrate(depth, T, rgb=np.broadcast_to(p, (H, W, 3)))
    ref.integrate(depth, T, rgb=np.broadcast_to(q, (H, W, 3)))
    keys, s, wt, _, _ = ref.export()
    _, col = ref.export_rgb()
    want = (p.astype(int) + q.astype(int)) // 2
    assert np.all(col[wt == 2] == want) and (wt == 2).sum() > 100000
    assert np.all(col[wt == 0] == 0)
    xyz, cubes, rgb = ref.mc(with_rgb=True)
    assert len(xyz) > 1000
    assert np.all(rgb.reshape(-1, 3) == want)


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_colour_weights_follow_the_storage_weight(oracle, wl, storage):
    """With a weight cap of 2 the colour becomes an exponential average: after frames of
    colours a, b, c (all observed) a voxel holds floor((floor((a + b) / 2) * 2 + c) / 3)."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    H, W = w["height"], w["width"]
    T = look_down_z()
    depth = plane_depth(w, 1000)
    cols = [np.array([100, 0, 255], np.uint8), np.array([201, 50, 0], np.uint8), np.array([0, 90, 3], np.uint8)]
    ref = oracle.Oracle(w, storage=storage, weight_cap=2, color=True)
    for c in cols:
        ref.integrate(depth, T, rgb=np.broadcast_to(c, (H, W, 3)))
    keys, s, wt, _, _ = ref.export()
    _, col = ref.export_rgb()
    a, b, c = (x.astype(int) for x in cols)
    ab = (a + b) // 2
    want = (ab * 2 + c) // 3
    m = wt == 2
    assert m.sum() > 100000 and np.all(col[m] == want)
