"""Pins the oracle's carving mode (every allocated voxel a frame observes is updated, as the
reference's whole-ray block selection integrates every existing block along the pixel rays
inside its frustum, src/tsdf.cu:2128-2177, :2056-2064).

Closed form: a plane at 1.0 m allocates its band's blocks; a second frame of a plane at 1.5 m
from the same pose sees those voxels in free space, so each voxel the first frame observed
takes the mean with dist = 1 and one more weight; without carving they keep their values.
Brute force: the per-block bound that skips blocks no voxel of which can be observed drops
nothing (the same volume as testing every voxel of every block each frame).
"""
import numpy as np

from conftest import look_down_z, plane_depth


def _two_planes(oracle, wl, carve):
    w = dict(wl.WORKLOADS["room_small_1cm"])
    ref = oracle.Oracle(w, carve=carve)
    T = look_down_z()
    ref.integrate(plane_depth(w, 1000), T)
    k1, s1, w1, s64_1, w64_1 = ref.export()
    ref.integrate(plane_depth(w, 1500), T)
    k2, s2, w2, s64_2, w64_2 = ref.export()
    return w, (k1, s1, w1, s64_1, w64_1), (k2, s2, w2, s64_2, w64_2)


def _lattice_z(keys, vox):
    """World z (= camera z for the identity pose) of every voxel, (N, 512)."""
    i = np.arange(512)
    return (keys[:, 2:3] * 8 + (i & 7)[None, :]).astype(np.float64) * vox


def test_carving_moves_free_space_toward_one(oracle, wl):
    w, (k1, s1, w1, _, _), (k2, s2, w2, s64, w64) = _two_planes(oracle, wl, carve=1)
    # the FP64 shadow takes the parameters as their float32 values
    f = lambda x: float(np.float32(x))
    mu, vox, d1 = f(w["trunc"]), f(w["vox_size"]), 1000 * f(0.001)
    # the first frame's blocks are all still there, in the same sorted positions
    idx = np.searchsorted(k2.view([("", np.int32)] * 3).ravel(), k1.view([("", np.int32)] * 3).ravel())
    assert np.array_equal(k2[idx], k1)
    s2, w2, s64, w64 = s2[idx], w2[idx], s64[idx], w64[idx]
    obs1 = w1 > 0
    assert obs1.sum() > 100_000
    # observed twice: the running mean of the first distance and 1, weight 2
    assert np.all(w2[obs1] == 2)
    assert np.array_equal(s2[obs1], (s1[obs1] + np.float32(1)) / np.float32(2))
    z = _lattice_z(k1, vox)
    first = np.minimum(1.0, (d1 - z) / mu)
    assert np.abs(s64[obs1] - (first[obs1] + 1.0) / 2.0).max() < 1e-12
    assert np.all(w64[obs1] == 2)
    # voxels the first frame did not observe but the second does (in the image, in front of
    # 1.5 m): first observation, dist = 1
    obs2 = (~obs1) & (w2 > 0)
    assert obs2.sum() > 0 and np.all(w2[obs2] == 1) and np.all(s2[obs2] == 1)


def test_band_only_leaves_free_space_unchanged(oracle, wl):
    w, (k1, s1, w1, _, _), (k2, s2, w2, _, _) = _two_planes(oracle, wl, carve=0)
    idx = np.searchsorted(k2.view([("", np.int32)] * 3).ravel(), k1.view([("", np.int32)] * 3).ravel())
    assert np.array_equal(k2[idx], k1)
    assert np.array_equal(s2[idx].view(np.uint32), s1.view(np.uint32))
    assert np.array_equal(w2[idx], w1)


def test_carving_bound_drops_nothing(oracle, wl, small_frames):
    """The per-block bound against testing every voxel of every block: bit-identical volumes
    over the 40 frames of the small room (the orbit inside the room leaves blocks behind the camera,
    beside the image and straddling the camera plane)."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = small_frames
    a = oracle.Oracle(w, carve=1)
    b = oracle.Oracle(w, carve=2)
    base = oracle.Oracle(w, carve=0)
    for k in range(len(depth)):
        a.integrate(depth[k], poses[k])
        b.integrate(depth[k], poses[k])
        base.integrate(depth[k], poses[k])
    ka, sa, wa, s64a, w64a = a.export()
    kb, sb, wb, s64b, w64b = b.export()
    assert np.array_equal(ka, kb)
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(wa, wb)
    assert np.array_equal(s64a, s64b) and np.array_equal(w64a, w64b)
    # carving changes the volume: more observations, never fewer, on the same block set
    k0, s0, w0, _, _ = base.export()
    assert np.array_equal(k0, ka)
    assert np.all(wa >= w0) and (wa > w0).sum() > 10_000
