"""GPU parity of carving (vh_config.carve = 1): each frame updates every allocated voxel it
observes, not only the blocks its truncation bands cross, so free space in front of the
surface moves toward +1 (the reference integrates every existing block along its pixel rays,
src/tsdf.cu:2128-2177). The GPU finds those blocks with a conservative per-block visibility
and depth test in a pass after each batch's allocation; the oracle tests every allocated block.

Bars: block set, voxel values (bit patterns) and weights identical to the oracle's carving
mode on every input path (host frames, batched device sequences, profiled launches, tracked
frames) and voxel format; meshes identical; a block allocated by one frame is carved by the
next (two planes, closed form in the oracle's pins).
"""
import numpy as np
import pytest

from conftest import look_down_z, plane_depth
from test_gpu_integration import assert_same_mesh, assert_same_volume

pytestmark = pytest.mark.gpu


def _seq(depth, poses):
    import torch

    return (torch.from_numpy(np.ascontiguousarray(depth).view(np.int16)).cuda(),
            torch.from_numpy(np.ascontiguousarray(poses, np.float32).reshape(len(poses), 16)).cuda())


@pytest.mark.parametrize("path", ["host", "sequence"])
def test_block_allocated_by_one_frame_is_carved_by_the_next(vh, oracle, wl, path):
    """A plane at 1.0 m, then a plane at 1.5 m from the same pose: with carving the first
    plane's blocks are observed in free space by the second frame (weight + 1, value toward
    +1), equal to the oracle; without it they keep their values. On the sequence path both
    frames are in one batch, so the block born in frame 0 is carved by frame 1 of its batch."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth = np.stack([plane_depth(w, 1000), plane_depth(w, 1500)])
    poses = np.stack([look_down_z(), look_down_z()])
    print(f"two planes, {path} path", flush=True)
    vols = {}
    for carve in (0, 1):
        vol = vh.Volume(vh.config_from_workload(w, carve=carve))
        if path == "host":
            for k in range(2):
                vol.integrate(depth[k], poses[k])
        else:
            d_depth, d_pose = _seq(depth, poses)
            vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, 2)
            vol.integrate_next(2)
            vol.sync()
        ref = oracle.Oracle(w, carve=carve)
        for k in range(2):
            ref.integrate(depth[k], poses[k])
        keys, s, wt, _ = assert_same_volume(oracle, vol, ref)
        st = vol.stats()
        assert (st["carved_total"] > 0) == (carve == 1), st["carved_total"]
        vols[carve] = (keys, s, wt)
        vol.close()
    (k0, s0, w0), (k1, s1, w1) = vols[0], vols[1]
    assert np.array_equal(k0, k1)
    near = (k0[:, 2] * 8 * 0.01 < 1.2)  # the first plane's blocks
    assert near.sum() > 100
    assert np.all(w1[near] >= w0[near]) and (w1[near] > w0[near]).sum() > 100_000
    assert np.all(s1[near][w0[near] > 0] >= s0[near][w0[near] > 0])


@pytest.mark.parametrize("variant", ["f32", "f16", "f16r", "f32_colour"])
def test_room_small_carving_sequence_matches_oracle(vh, oracle, wl, small_frames, variant):
    import torch

    w = dict(wl.WORKLOADS["room_small_1cm"])
    over = {"f32": {}, "f16": dict(storage=1, weight_cap=255), "f16r": dict(storage=2, weight_cap=255),
            "f32_colour": dict(color=1)}[variant]
    depth, poses = small_frames
    n = len(depth)
    print(f"room_small_1cm: {n} sequence frames, batch 8, carving, {variant}", flush=True)
    cfg = vh.config_from_workload(w, carve=1, **over)
    vol = vh.Volume(cfg)
    d_depth, d_pose = _seq(depth, poses)
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    colour = "color" in over
    if colour:
        scene = vh.Scene(w["scene"], wl.SEED)
        rgb = np.stack([scene.rgb(k, poses[k]) for k in range(n)])
        vol.bind_sequence_rgb(torch.from_numpy(rgb).cuda())
    vol.integrate_next(n)
    vol.sync()
    ref = oracle.Oracle(w, storage=cfg.storage, weight_cap=cfg.weight_cap, color=colour, carve=1)
    touched = 0
    for k in range(n):
        touched += ref.integrate(depth[k], poses[k], rgb=rgb[k] if colour else None)
    st = vol.stats()
    assert st["frames"] == n and st["touched_total"] == touched
    assert st["carved_total"] > 0
    assert_same_volume(oracle, vol, ref)
    if colour:
        keys, _, _ = vol.export_blocks()
        order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
        _, rrgb = ref.export_rgb()
        assert np.array_equal(vol.export_block_colors()[order], rrgb)
    assert_same_mesh(vol, ref)
    vol.close()


def test_room_small_carving_host_and_profiled_paths(vh, oracle, wl, small_frames):
    """Host frames (one allocation, carve and integration per frame), then the rest of the
    sequence in profiling mode (eager launches with events around the carve pass)."""
    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = small_frames
    print("room_small_1cm: 20 host frames, then 20 profiled sequence frames, carving", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, carve=1))
    ref = oracle.Oracle(w, carve=1)
    for k in range(20):
        vol.integrate(depth[k], poses[k])
        assert vol.stats()["touched_last"] == ref.integrate(depth[k], poses[k])
    assert_same_volume(oracle, vol, ref)
    d_depth, d_pose = _seq(depth[20:], poses[20:])
    vol.set_profiling(True)
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, 20)
    vol.integrate_next(20)
    vol.sync()
    st = vol.stats()
    vol.set_profiling(False)
    assert st["profiled_frames"] == 20 and st["carve_ms"] > 0
    for k in range(20, 40):
        ref.integrate(depth[k], poses[k])
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    vol.close()
