"""GPU parity of allocation, integration and marching cubes against the CPU oracle.

Bar: the allocated block set, per-frame touched counts, FP32 (and FP16-storage) voxel
values and the marching-cubes triangles are bit-identical to the oracle's twin; the FP32
values stay within 5e-5 of the oracle's FP64 shadow (normalised TSDF units).
"""
import numpy as np
import pytest

from conftest import look_down_z, plane_depth

pytestmark = pytest.mark.gpu


def run_pair(vh, oracle, w, depth, poses, **over):
    cfg = vh.config_from_workload(w, **over)
    vol = vh.Volume(cfg)
    ref = oracle.Oracle(w, storage=cfg.storage, weight_cap=cfg.weight_cap)
    touched = []
    for k in range(len(depth)):
        vol.integrate(depth[k], poses[k])
        n_ref = ref.integrate(depth[k], poses[k])
        touched.append((vol.stats()["touched_last"], n_ref))
    return vol, ref, touched


def gpu_sorted(oracle, vol):
    keys, s, wt = vol.export_blocks()
    return oracle.sort_blocks(keys, s, wt)


def assert_same_volume(oracle, vol, ref):
    keys, s, wt = gpu_sorted(oracle, vol)
    rk, rs, rw, r64, rw64 = ref.export()
    assert keys.shape == rk.shape and np.array_equal(keys, rk), "allocated block sets differ"
    assert np.array_equal(wt, rw), "weights differ"
    bad = s.view(np.uint32) != rs.view(np.uint32)
    assert not bad.any(), f"{bad.sum()} tsdf values differ, max {np.abs(s - rs).max()}"
    return keys, s, wt, r64


def assert_same_mesh(vol, ref):
    xyz, cubes = vol.extract_mesh()
    rxyz, rcubes = ref.mc()
    assert len(xyz) == len(rxyz)
    o = np.lexsort(cubes.T[::-1])
    ro = np.lexsort(rcubes.T[::-1])
    assert np.array_equal(cubes[o], rcubes[ro]), "triangle topology differs"
    assert np.array_equal(xyz[o].view(np.uint32), rxyz[ro].view(np.uint32)), "vertex positions differ"
    return len(xyz)


def test_room_small_fp32_bit_exact(vh, oracle, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    print(f"room_small_1cm: {len(depth)} frames FP32", flush=True)
    vol, ref, touched = run_pair(vh, oracle, w, depth, poses)
    assert all(a == b for a, b in touched), touched
    keys, s, wt, r64 = assert_same_volume(oracle, vol, ref)
    m = wt > 0
    err = np.abs(s[m] - r64[m])
    print(f"blocks {len(keys)}, FP32 vs FP64 max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() < 5e-5
    n = assert_same_mesh(vol, ref)
    assert n > 1000
    st = vol.stats()
    assert st["frames"] == len(depth) and st["blocks_allocated"] == len(keys)
    assert st["touched_total"] == sum(a for a, _ in touched)
    vol.close()


# binary16 running means accumulate one rounding per store (~ulp/2 = 2.4e-4 near |t| = 1),
# so plain FP16 drifts to a few 1e-3 after tens of frames; the residual byte of F16R carries
# that rounding forward and keeps the error near 1e-5.
@pytest.mark.parametrize("storage,bound", [(1, 4e-3), (2, 5e-5)], ids=["f16", "f16r"])
def test_room_small_16bit_storage_bit_exact(vh, oracle, wl, small_frames, storage, bound):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    print(f"room_small_1cm: 40 frames storage {storage}, weight cap 255", flush=True)
    vol, ref, touched = run_pair(vh, oracle, w, depth, poses, storage=storage, weight_cap=255)
    assert all(a == b for a, b in touched)
    keys, s, wt, r64 = assert_same_volume(oracle, vol, ref)
    m = wt > 0
    err = np.abs(s[m] - r64[m])
    print(f"storage {storage} vs FP64 max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e}")
    assert err.max() < bound
    assert_same_mesh(vol, ref)
    vol.close()


def test_device_sequence_graph_matches_host_path(vh, oracle, wl, small_frames):
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 12
    print("room_small_1cm: 12 frames through bind_sequence + graph launches", flush=True)
    d_depth = torch.from_numpy(depth[:n].view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    torch.cuda.synchronize()
    vol = vh.Volume(vh.config_from_workload(w))
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    vol.integrate_next(n)
    vol.sync()
    ref = oracle.Oracle(w)
    for k in range(n):
        ref.integrate(depth[k], poses[k])
    assert vol.stats()["frames"] == n
    assert_same_volume(oracle, vol, ref)
    # profiling mode (eager launches with events) continues the same sequence identically
    vol2 = vh.Volume(vh.config_from_workload(w))
    vol2.set_profiling(True)
    vol2.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    vol2.integrate_next(n)
    vol2.sync()
    st = vol2.stats()
    assert st["profiled_frames"] == n and st["integrate_ms"] > 0 and st["alloc_ms"] > 0
    assert_same_volume(oracle, vol2, ref)
    vol.close()
    vol2.close()


def test_sequence_chunks_overlap_allocation_and_integration(vh, oracle, wl, small_frames):
    """Bound sequences run one launch per frame that integrates frame f and allocates frame
    f+1; chunked calls (1, 5, 0, then 6 frames, a host-staged frame in between) must give the
    same volume, statistics and mesh as the oracle frame by frame."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 12
    print("room_small_1cm: 12 sequence frames in chunks 1, 5, 0, 6 plus one staged frame", flush=True)
    d_depth = torch.from_numpy(depth[:n].view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    touched = []
    done = 0
    for chunk in (1, 5, 0):
        vol.integrate_next(chunk)
        vol.sync()
        for k in range(done, done + chunk):
            touched.append(ref.integrate(depth[k], poses[k]))
        done += chunk
        if chunk:
            assert vol.stats()["touched_last"] == touched[-1]
    # a host-staged frame between chunks (frame 20 of the same camera path)
    vol.integrate(depth[20], poses[20])
    touched.append(ref.integrate(depth[20], poses[20]))
    vol.integrate_next(6)
    vol.sync()
    for k in range(done, n):
        touched.append(ref.integrate(depth[k], poses[k]))
    st = vol.stats()
    assert st["frames"] == n + 1 and st["touched_last"] == touched[-1] and st["touched_total"] == sum(touched)
    assert st["new_total"] == ref.num_blocks()
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    vol.close()


def test_async_host_frames_match_oracle(vh, oracle, wl, small_frames):
    """vh_integrate_async: frames copied on a second stream into a 4-slot device ring (more
    frames than slots, so slots are reused), page-locked and pageable sources, float depth,
    and a synchronous staged frame in between; volume and mesh equal to the oracle's."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    print("room_small_1cm: 14 host frames through the asynchronous copy ring", flush=True)
    pinned = torch.from_numpy(depth[:10].view(np.int16).copy()).pin_memory()
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    for k in range(10):
        vol.integrate_async(pinned[k], poses[k])
    vol.integrate(depth[10], poses[10])  # staged, synchronous, between queued frames
    for k in (11, 12):
        vol.integrate_async(depth[k], poses[k])  # pageable numpy
    vol.integrate_async(depth[13].astype(np.float32) * np.float32(0.001), poses[13])
    vol.sync()
    for k in range(14):
        ref.integrate(depth[k], poses[k])
    assert vol.stats()["frames"] == 14
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    # after a reset the ring restarts in step with the device
    vol.reset()
    for k in range(6):
        vol.integrate_async(pinned[k], poses[k])
    vol.sync()
    ref2 = oracle.Oracle(w)
    for k in range(6):
        ref2.integrate(depth[k], poses[k])
    assert_same_volume(oracle, vol, ref2)
    vol.close()


def test_async_frames_may_change_depth_format(vh, oracle, wl, small_frames):
    """Queued frames of both depth formats in any order (the ring's slots have the element
    size of the frames in flight; a format change waits for them): equal to the oracle."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    print("room_small_1cm: 16 asynchronous frames alternating uint16 / float32 depth", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    order = [0, 0, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, 1, 1, 0, 1]
    pinned = torch.from_numpy(depth[:16].view(np.int16).copy()).pin_memory()
    as_float = torch.from_numpy(depth[:16].astype(np.float32) * np.float32(0.001)).pin_memory()
    for k, fmt in enumerate(order):
        vol.integrate_async(as_float[k] if fmt else pinned[k], poses[k])
    vol.sync()
    for k in range(16):
        ref.integrate(depth[k], poses[k])
    assert vol.stats()["frames"] == 16
    assert_same_volume(oracle, vol, ref)
    vol.close()


def test_float_depth_input_matches_u16(vh, oracle, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w))
    for k in range(4):
        vol.integrate(depth[k].astype(np.float32) * np.float32(0.001), poses[k])
    ref = oracle.Oracle(w)
    for k in range(4):
        ref.integrate(depth[k], poses[k])
    assert_same_volume(oracle, vol, ref)
    vol.close()


def test_non_finite_and_non_positive_float_depth_is_no_measurement(vh, oracle, wl, small_frames):
    """Float depth pixels that are NaN, +-inf, negative, -0 or beyond max_depth are no
    measurement, exactly like 0: the volume equals the oracle's given 0 at those pixels, on
    the host path and through a bound device sequence (batched graph launches). The
    reference's rejection test `d <= 0 || d > max_depth` (src/tsdf.cu:713) would let NaN
    through; its inputs (16-bit PNG depth) never hold NaN."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 6
    rng = np.random.default_rng(713)
    specials = np.array([np.nan, np.inf, -np.inf, -0.5, -0.0, w["max_depth"] + 1.0], np.float32)
    fd = depth[:n].astype(np.float32) * np.float32(0.001)
    u = depth[:n].copy()
    pick = rng.random(fd.shape) < 0.12
    fd[pick] = specials[rng.integers(0, len(specials), int(pick.sum()))]
    u[pick] = 0
    print(f"room_small_1cm: {n} float frames, {int(pick.sum())} special pixels", flush=True)
    ref = oracle.Oracle(w)
    for k in range(n):
        ref.integrate(u[k], poses[k])
    vol = vh.Volume(vh.config_from_workload(w))
    for k in range(n):
        vol.integrate(fd[k], poses[k])
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    vol.close()
    d_depth = torch.from_numpy(fd).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    torch.cuda.synchronize()
    seq = vh.Volume(vh.config_from_workload(w))
    seq.bind_sequence(d_depth, vh.DEPTH_F32, 1.0, d_pose, n)
    seq.integrate_next(n)
    seq.sync()
    assert_same_volume(oracle, seq, ref)
    seq.close()


def test_odd_image_size_partial_tiles(vh, oracle, wl, small_frames):
    """A 203 x 151 window of the 640 x 480 frames (principal point shifted by the window's
    origin): neither dimension is a multiple of the allocation tiles (16 x 16 for single
    frames, 16 x 64 in batched launches), so edge tiles are partial in both directions.
    Single frames and a bound sequence in batches of 8 equal the oracle bit for bit."""
    import torch

    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = small_frames
    u0, v0, W, H = 211, 167, 203, 151
    w.update(width=W, height=H, cx=w["cx"] - u0, cy=w["cy"] - v0)
    n = 12
    crop = np.ascontiguousarray(depth[:n, v0:v0 + H, u0:u0 + W])
    print(f"room_small_1cm window {W}x{H} at ({u0}, {v0}): {n} frames", flush=True)
    ref = oracle.Oracle(w)
    vol = vh.Volume(vh.config_from_workload(w))
    for k in range(n):
        ref.integrate(crop[k], poses[k])
        vol.integrate(crop[k], poses[k])
    keys, _, _, _ = assert_same_volume(oracle, vol, ref)
    assert len(keys) > 100  # the window sees one wall: 113 blocks of 8 cm
    assert_same_mesh(vol, ref)
    vol.close()
    d_depth = torch.from_numpy(crop.view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    torch.cuda.synchronize()
    seq = vh.Volume(vh.config_from_workload(w, seq_batch=8))
    seq.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    seq.integrate_next(n)
    seq.sync()
    assert_same_volume(oracle, seq, ref)
    seq.close()


def test_far_from_origin_near_the_key_range(vh, oracle, wl, small_frames):
    """Every pose shifted by (83, -83, 41) km: block coordinates near +-1.04e6, close to the
    allocation's limit of 1,048,000 blocks per axis (21-bit keys), lattice indices near
    8.3e6. FP32 spacing there is ~8 mm, so the geometry is coarse, but the GPU and the oracle
    take the same float inputs: block set, voxels and mesh equal bit for bit."""
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 8
    far = poses[:n].copy()
    far[:, 0, 3] += np.float32(83000.0)
    far[:, 1, 3] -= np.float32(83000.0)
    far[:, 2, 3] += np.float32(41000.0)
    print("room_small_1cm: 8 frames at (83, -83, 41) km", flush=True)
    vol, ref, touched = run_pair(vh, oracle, w, depth[:n], far)
    assert all(a == b for a, b in touched)
    keys, _, _, _ = assert_same_volume(oracle, vol, ref)
    assert len(keys) > 500  # 574: the first frames face one wall
    assert keys[:, 0].min() > 1_030_000 and keys[:, 1].max() < -1_030_000
    assert assert_same_mesh(vol, ref) > 1000
    vol.close()


def test_pitched_depth_rows(vh, oracle, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    padded = np.zeros((480, 700), np.uint16)
    padded[:, :640] = depth[0]
    vol = vh.Volume(vh.config_from_workload(w))
    vol.integrate(padded[:, :640], poses[0])  # non-contiguous view: pitch 1400 bytes
    ref = oracle.Oracle(w)
    ref.integrate(depth[0], poses[0])
    assert_same_volume(oracle, vol, ref)
    vol.close()


def test_plane_closed_form_on_gpu(vh, oracle, wl):
    w = wl.WORKLOADS["room_small_1cm"]
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    T = look_down_z(0.013, -0.021, 0.0)
    for D in (1500, 1500, 1512):
        vol.integrate(plane_depth(w, D), T)
        ref.integrate(plane_depth(w, D), T)
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    vol.close()


def test_empty_and_out_of_range_frames(vh, wl):
    w = wl.WORKLOADS["room_small_1cm"]
    vol = vh.Volume(vh.config_from_workload(w))
    vol.integrate(plane_depth(w, 0), look_down_z())
    vol.integrate(plane_depth(w, 6000), look_down_z())  # beyond max_depth
    st = vol.stats()
    assert st["blocks_allocated"] == 0 and st["touched_last"] == 0 and st["frames"] == 2
    xyz, cubes = vol.extract_mesh()
    assert len(xyz) == 0
    with pytest.raises(vh.VhError) as e:
        vol.integrate(plane_depth(w, 1500), np.full((4, 4), np.nan, np.float32))
    assert e.value.code == vh.VH_ERR_INVALID_ARG
    far = look_down_z(3.0e5, 0.0, 0.0)  # 3e5 m / 8 cm blocks: beyond the 21-bit key range
    with pytest.raises(vh.VhError) as e:
        vol.integrate(plane_depth(w, 1500), far)
    assert e.value.code == vh.VH_ERR_RANGE
    vol.close()


def test_pool_full_is_reported(vh, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w, pool_blocks=64))
    with pytest.raises(vh.VhError) as e:
        vol.integrate(depth[0], poses[0])
    assert e.value.code == vh.VH_ERR_POOL_FULL
    st = vol.stats()
    assert st["blocks_allocated"] == 64
    vol.close()


def test_hash_full_is_reported(vh, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w, log2_buckets=4))  # 64 slots in total
    with pytest.raises(vh.VhError) as e:
        vol.integrate(depth[0], poses[0])
    assert e.value.code == vh.VH_ERR_HASH_FULL
    vol.close()


def test_reset_empties_the_volume(vh, oracle, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w))
    for k in range(3):
        vol.integrate(depth[k], poses[k])
    vol.reset()
    assert vol.stats()["blocks_allocated"] == 0
    vol.integrate(depth[5], poses[5])
    ref = oracle.Oracle(w)
    ref.integrate(depth[5], poses[5])
    assert_same_volume(oracle, vol, ref)
    vol.close()


@pytest.mark.parametrize("batch", [1, 3, 8])
def test_batched_sequences_equal_frame_by_frame(vh, oracle, wl, small_frames, batch):
    """Bound sequences integrate `batch` frames per launch (each block read and written once
    per batch, updated by every frame of the batch that touched it, in frame order): the
    volume, the per-frame statistics and the mesh equal the oracle's frame-by-frame result,
    with calls of 5 and 7 frames (partial batches) and a profiled call; blocks_integrated is
    the sum over batches of the union of the frames' touched sets."""
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 16
    print(f"room_small_1cm: {n} sequence frames, batch {batch}, calls 5 + 7 + 4 (profiled)", flush=True)
    d_depth = torch.from_numpy(depth[:n].view(np.int16).copy()).cuda()
    d_pose = torch.from_numpy(poses[:n].copy()).cuda()
    vol = vh.Volume(vh.config_from_workload(w, seq_batch=batch))
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, n)
    ref = oracle.Oracle(w)
    touched = []
    sets = []

    def ref_frames(k0, k1):
        for k in range(k0, k1):
            touched.append(ref.integrate(depth[k], poses[k]))
            sets.append({tuple(x) for x in ref.touched()})

    vol.integrate_next(5)
    vol.integrate_next(7)
    vol.sync()
    ref_frames(0, 12)
    st = vol.stats()
    assert st["frames"] == 12 and st["touched_last"] == touched[-1] and st["touched_total"] == sum(touched)
    assert_same_volume(oracle, vol, ref)
    vol.set_profiling(True)
    vol.integrate_next(4)
    vol.sync()
    ref_frames(12, 16)
    st = vol.stats()
    vol.set_profiling(False)
    assert st["profiled_frames"] == 4 and st["profiled_launches"] == -(-4 // batch)
    assert st["touched_total"] == sum(touched) and st["touched_last"] == touched[-1]
    # batches: calls of 5, 7 and 4 frames, each cut into runs of `batch`
    expect = 0
    f = 0
    for call in (5, 7, 4):
        for b0 in range(0, call, batch):
            u = set().union(*sets[f + b0:f + min(call, b0 + batch)])
            expect += len(u)
        f += call
    assert st["blocks_integrated"] == expect, (st["blocks_integrated"], expect)
    assert_same_volume(oracle, vol, ref)
    assert_same_mesh(vol, ref)
    vol.close()
