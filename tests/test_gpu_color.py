"""GPU parity of the colour channel against the CPU oracle.

Bar: bit-identical voxel colours, TSDF values and marching-cubes vertex colours for every
storage format and every input path (host frames, the asynchronous copy ring, bound
device sequences, tracked sequences), with a textured RGB image that changes per frame.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def texture(H, W, f):
    u = np.arange(W)[None, :]
    v = np.arange(H)[:, None]
    r = (u * 7 + f * 13) & 255
    g = (v * 5 + 3 * f) & 255
    b = ((u ^ v) + 11 * f) & 255
    return np.ascontiguousarray(np.stack(np.broadcast_arrays(r, g, b), axis=-1).astype(np.uint8))


def colour_frames(w, n):
    return np.stack([texture(w["height"], w["width"], f) for f in range(n)])


def assert_same_colour_volume(oracle, vol, ref):
    keys, s, wt = vol.export_blocks()
    rgb = vol.export_block_colors()
    keys, s, wt, rgb = oracle.sort_blocks(keys, s, wt, rgb)
    rk, rs, rw, _, _ = ref.export()
    rk2, rrgb = ref.export_rgb()
    assert np.array_equal(keys, rk) and np.array_equal(rk2, rk), "allocated block sets differ"
    assert np.array_equal(wt, rw), "weights differ"
    assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), "tsdf values differ"
    bad = (rgb != rrgb).any(-1)
    assert not bad.any(), f"{bad.sum()} voxel colours differ"
    return keys, wt, rgb


def assert_same_colour_mesh(vol, ref):
    xyz, cubes, rgb = vol.extract_mesh(with_rgb=True)
    rxyz, rcubes, rrgb = ref.mc(with_rgb=True)
    assert rgb is not None and len(xyz) == len(rxyz)
    o = np.lexsort(cubes.T[::-1])
    ro = np.lexsort(rcubes.T[::-1])
    assert np.array_equal(cubes[o], rcubes[ro]), "triangle topology differs"
    assert np.array_equal(xyz[o].view(np.uint32), rxyz[ro].view(np.uint32)), "vertex positions differ"
    assert np.array_equal(rgb[o], rrgb[ro]), f"{(rgb[o] != rrgb[ro]).sum()} vertex colour bytes differ"
    return len(xyz), rgb


@pytest.mark.parametrize("storage,cap", [(0, 0), (1, 255), (2, 255)], ids=["f32", "f16", "f16r"])
def test_host_frames_colour_bit_exact(vh, oracle, wl, small_frames, storage, cap):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 16
    rgb = colour_frames(w, n)
    print(f"room_small_1cm colour: {n} host frames, storage {storage}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, storage=storage, weight_cap=cap, color=1))
    ref = oracle.Oracle(w, storage=storage, weight_cap=cap, color=True)
    for k in range(n):
        vol.integrate(depth[k], poses[k], rgb=rgb[k])
        ref.integrate(depth[k], poses[k], rgb=rgb[k])
    keys, wt, col = assert_same_colour_volume(oracle, vol, ref)
    assert (wt > 1).sum() > 100000 and len(np.unique(col[wt > 0].reshape(-1, 3), axis=0)) > 1000
    ntri, mrgb = assert_same_colour_mesh(vol, ref)
    assert ntri > 1000 and mrgb.max() > 0
    vol.close()


def test_async_sequence_and_tracked_colour_paths(vh, oracle, wl, small_frames):
    """The same 12 frames through the asynchronous ring (pinned RGB), a bound device
    sequence, and a tracked sequence with zero solver iterations (which integrates at the
    prior poses): each volume equals the oracle's."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 12
    scene = vh.Scene(w["scene"], wl.SEED)
    rgb = np.stack([scene.rgb(k, poses[k]) for k in range(n)])  # the scene's colours
    ref = oracle.Oracle(w, color=True)
    for k in range(n):
        ref.integrate(depth[k], poses[k], rgb=rgb[k])
    cfg = vh.config_from_workload(w, color=1)

    print("colour: asynchronous ring", flush=True)
    pinned_d = torch.from_numpy(depth[:n].view(np.int16).copy()).pin_memory()
    pinned_c = torch.from_numpy(rgb.copy()).pin_memory()
    vol = vh.Volume(cfg)
    for k in range(n):
        vol.integrate_async(pinned_d[k], poses[k], rgb=pinned_c[k])
    vol.sync()
    assert_same_colour_volume(oracle, vol, ref)
    vol.close()

    print("colour: bound device sequence", flush=True)
    d_depth = torch.from_numpy(depth[:n].view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    d_rgb = torch.from_numpy(rgb.copy()).cuda()
    torch.cuda.synchronize()
    vol = vh.Volume(cfg)
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    with pytest.raises(vh.VhError):
        vol.integrate_next(1)  # a colour volume needs the sequence's RGB
    vol.bind_sequence_rgb(d_rgb)
    vol.integrate_next(5)
    vol.integrate_next(n - 5)
    vol.sync()
    assert_same_colour_volume(oracle, vol, ref)
    assert_same_colour_mesh(vol, ref)
    vol.close()

    print("colour: tracked sequence, 0 iterations", flush=True)
    d_out = torch.zeros_like(d_pose)
    vol = vh.Volume(cfg)
    vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, n, d_pose, d_out, gn_iters=0, point_stride=8)
    vol.bind_sequence_rgb(d_rgb)
    vol.integrate_next(n)
    vol.sync()
    assert np.array_equal(d_out.cpu().numpy(), poses[:n])
    assert_same_colour_volume(oracle, vol, ref)
    vol.close()


def test_colour_odd_image_size_batched(vh, oracle, wl, small_frames):
    """A 203 x 151 window (principal point shifted by its origin) of depth and texture: the
    allocation tiles, which also pack each pixel's RGB for the integration, are partial in
    both directions. Host frames and a bound sequence in batches of 8 equal the oracle."""
    import torch

    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = small_frames
    u0, v0, W, H = 211, 167, 203, 151
    w.update(width=W, height=H, cx=w["cx"] - u0, cy=w["cy"] - v0)
    n = 10
    crop = np.ascontiguousarray(depth[:n, v0:v0 + H, u0:u0 + W])
    rgb = np.ascontiguousarray(np.stack([texture(480, 640, f) for f in range(n)])[:, v0:v0 + H, u0:u0 + W])
    print(f"colour window {W}x{H}: {n} frames, host and batched sequence", flush=True)
    ref = oracle.Oracle(w, color=True)
    for k in range(n):
        ref.integrate(crop[k], poses[k], rgb=rgb[k])
    cfg = vh.config_from_workload(w, color=1, seq_batch=8)
    vol = vh.Volume(cfg)
    for k in range(n):
        vol.integrate(crop[k], poses[k], rgb=rgb[k])
    assert_same_colour_volume(oracle, vol, ref)
    assert_same_colour_mesh(vol, ref)
    vol.close()
    d_depth = torch.from_numpy(crop.view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    d_rgb = torch.from_numpy(rgb.copy()).cuda()
    torch.cuda.synchronize()
    seq = vh.Volume(cfg)
    seq.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    seq.bind_sequence_rgb(d_rgb)
    seq.integrate_next(n)
    seq.sync()
    assert_same_colour_volume(oracle, seq, ref)
    assert_same_colour_mesh(seq, ref)
    seq.close()


def test_colour_argument_errors(vh, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w, color=1))
    with pytest.raises(vh.VhError) as e:
        vol.integrate(depth[0], poses[0])  # no RGB
    assert e.value.code == vh.VH_ERR_INVALID_ARG
    assert vol.stats()["frames"] == 0
    vol.close()
    plain = vh.Volume(vh.config_from_workload(w))
    plain.integrate(depth[0], poses[0], rgb=texture(w["height"], w["width"], 0))  # ignored
    import torch

    d = torch.zeros(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(vh.VhError):
        plain.bind_sequence_rgb(d)
    with pytest.raises(vh.VhError):
        plain.export_block_colors()
    xyz, cubes, rgb = plain.extract_mesh(with_rgb=True)
    assert rgb is None and len(xyz) > 0
    plain.close()
