"""Out-of-core streaming (vh_stream_out / vh_stream_in) and checkpoints (vh_save_volume /
vh_load_volume) on the GPU.

Bar: a sequence integrated while blocks stream out and back in around the moving camera ends
bit-identical to the oracle that never streamed (block set, TSDF bits, weights, colours,
triangle soup); blocks allocated again while their old copy was out merge by the documented
weighted mean, checked against a float32 numpy restatement; a checkpoint reloaded and
integrated further equals the uninterrupted sequence bit for bit, whether its blocks were
resident or in the host store; errors leave the volume unchanged.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_DEPTH = 2.0  # metres: the orbit in room_medium_5mm (8 x 6 m) then leaves blocks out of reach


def _ray_factor(w):
    """Longest pixel ray per unit depth (the farther image corner)."""
    dx = max(w["cx"], w["width"] - 1 - w["cx"]) / w["fx"]
    dy = max(w["cy"], w["height"] - 1 - w["cy"]) / w["fy"]
    return math.sqrt(1 + dx * dx + dy * dy)


def _sorted_volume(oracle, vol, color):
    keys, s, wt = vol.export_blocks()
    arrays = [s, wt] + ([vol.export_block_colors()] if color else [])
    return oracle.sort_blocks(keys, *arrays)


def _assert_same_as_oracle(oracle, vol, ref, color):
    got = _sorted_volume(oracle, vol, color)
    rk, rs, rw, _, _ = ref.export()
    keys, s, wt = got[:3]
    assert np.array_equal(keys, rk), "block sets differ"
    assert np.array_equal(s.view(np.uint32), rs.view(np.uint32)), "tsdf values differ"
    assert np.array_equal(wt, rw), "weights differ"
    if color:
        _, rrgb = ref.export_rgb()
        assert np.array_equal(got[3], rrgb), "colours differ"
    xyz, cubes = vol.extract_mesh()
    rxyz, rcubes = ref.mc()
    o, ro = np.lexsort(cubes.T[::-1]), np.lexsort(rcubes.T[::-1])
    assert np.array_equal(cubes[o], rcubes[ro]) and np.array_equal(xyz[o].view(np.uint32), rxyz[ro].view(np.uint32))


@pytest.mark.parametrize("storage,cap,color,carve", [(0, 0, False, 0), (2, 255, True, 0), (0, 0, False, 1)],
                         ids=["f32", "f16r_colour", "f32_carve"])
def test_streamed_sequence_equals_oracle(vh, oracle, wl, storage, cap, color, carve):
    """room_medium_5mm, every 5th of its 500 frames (the camera orbits 1.15 m around the room
    centre), depth up to 2 m, in groups of 5 frames: before each group, blocks farther than
    the group's reach + 0.25 m go out and blocks within its reach come in."""
    from test_gpu_color import texture

    w = dict(wl.WORKLOADS["room_medium_5mm"], max_depth=MAX_DEPTH)
    scene = vh.Scene(w["scene"], wl.SEED)
    frames = list(range(0, 500, 5))
    depth, poses = scene.frames(0, 500)
    group = 5
    print(f"streaming: room_medium_5mm, {len(frames)} frames, max depth {MAX_DEPTH} m, storage {storage}, "
          f"colour {color}, carve {carve}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, storage=storage, weight_cap=cap, color=int(color), carve=carve))
    ref = oracle.Oracle(w, storage=storage, weight_cap=cap, color=color, carve=carve)
    reach = (MAX_DEPTH + w["trunc"]) * _ray_factor(w) + 8 * w["vox_size"] * math.sqrt(3)
    moved_out = moved_in = 0
    peak_host = 0
    for g0 in range(0, len(frames), group):
        fr = frames[g0:g0 + group]
        cams = poses[fr, :3, 3].astype(np.float64)
        centre = cams.mean(0)
        r_in = reach + np.linalg.norm(cams - centre, axis=1).max() + 1e-3
        moved_out += vol.stream_out(centre, r_in + 0.25)
        moved_in += vol.stream_in(centre, r_in)
        peak_host = max(peak_host, vol.stats()["host_blocks"])
        for k in fr:
            rgb = texture(w["height"], w["width"], k) if color else None
            vol.integrate(depth[k], poses[k], rgb=rgb)
            ref.integrate(depth[k], poses[k], rgb=rgb)
    moved_in += vol.stream_in()  # everything back
    st = vol.stats()
    print(f"streamed out {moved_out}, in {moved_in}, host store peak {peak_host}, merged "
          f"{st['stream_merged_total']}, blocks {st['blocks_allocated']}", flush=True)
    assert moved_out > 5000 and moved_in == moved_out and peak_host > 5000
    assert st["host_blocks"] == 0 and st["stream_merged_total"] == 0
    assert st["streamed_out_total"] == moved_out and st["streamed_in_total"] == moved_in
    _assert_same_as_oracle(oracle, vol, ref, color)
    vol.close()


def test_stream_in_merges_blocks_allocated_again(vh, oracle, wl, small_frames):
    """Frames 0-4, all blocks out (radius 0), frames 5-9 allocate part of them again, then
    everything in: resident-only and store-only blocks come back unchanged, blocks in both
    merge per voxel as (s_r w_r + s_s w_s) / (w_r + w_s) with weights added and colours
    likewise, truncated (float32 numpy restatement); a side with weight 0 keeps the other."""
    from test_gpu_color import colour_frames

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    rgb = colour_frames(w, 10)
    print("streaming: merge of blocks allocated again, FP32 + colour", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, color=1))
    for k in range(5):
        vol.integrate(depth[k], poses[k], rgb=rgb[k])
    ka, sa, wa, ca = _sorted_volume(oracle, vol, True)
    n_out = vol.stream_out((0.0, 0.0, 0.0), 0.0)
    st = vol.stats()
    assert n_out == len(ka) and st["blocks_allocated"] == 0 and st["host_blocks"] == len(ka)
    assert len(vol.extract_mesh()[0]) == 0
    for k in range(5, 10):
        vol.integrate(depth[k], poses[k], rgb=rgb[k])
    kb, sb, wb, cb = _sorted_volume(oracle, vol, True)
    assert vol.stream_in() == len(ka)
    kf, sf, wf, cf = _sorted_volume(oracle, vol, True)
    st = vol.stats()
    # expected: the union of both block sets
    ia = {tuple(k): i for i, k in enumerate(ka.tolist())}
    ib = {tuple(k): i for i, k in enumerate(kb.tolist())}
    both = sorted(set(ia) & set(ib))
    assert len(both) > 300 and st["stream_merged_total"] == len(both)
    assert [tuple(k) for k in kf.tolist()] == sorted(set(ia) | set(ib))
    for i, k in enumerate(kf.tolist()):
        k = tuple(k)
        if k in ia and k in ib:
            continue
        s_, w_, c_ = (sa, wa, ca) if k in ia else (sb, wb, cb)
        j = ia[k] if k in ia else ib[k]
        assert np.array_equal(sf[i].view(np.uint32), s_[j].view(np.uint32)) and np.array_equal(wf[i], w_[j])
        assert np.array_equal(cf[i], c_[j])
    fi = np.array([i for i, k in enumerate(map(tuple, kf.tolist())) if k in ia and k in ib])
    ja = np.array([ia[tuple(k)] for k in kf[fi].tolist()])
    jb = np.array([ib[tuple(k)] for k in kf[fi].tolist()])
    sr, wr, cr = sb[jb], wb[jb], cb[jb].astype(np.float32)  # resident side: frames 5-9
    ss, ws, cs = sa[ja], wa[ja], ca[ja].astype(np.float32)  # staged side: frames 0-4
    wsum = (wr + ws).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        s_m = ((sr * wr).astype(np.float32) + (ss * ws).astype(np.float32)).astype(np.float32) / wsum
        c_m = ((cr * wr[..., None]).astype(np.float32) + (cs * ws[..., None]).astype(np.float32)) / wsum[..., None]
    c_m = np.where(np.isfinite(c_m), c_m, 0).astype(np.float32).astype(np.uint8)
    take_r, take_s = ~(ws > 0), ~(wr > 0) & (ws > 0)
    exp_s = np.where(take_r, sr, np.where(take_s, ss, s_m)).astype(np.float32)
    exp_w = np.where(take_r, wr, np.where(take_s, ws, wsum)).astype(np.float32)
    exp_c = np.where(take_r[..., None], cb[jb], np.where(take_s[..., None], ca[ja], c_m))
    assert ((wr > 0) & (ws > 0)).sum() > 10000  # voxels observed by both halves
    assert np.array_equal(sf[fi].view(np.uint32), exp_s.view(np.uint32)), "merged tsdf differs"
    assert np.array_equal(wf[fi], exp_w), "merged weights differ"
    assert np.array_equal(cf[fi], exp_c), "merged colours differ"
    vol.close()


@pytest.mark.parametrize("storage,cap,color", [(0, 0, False), (2, 255, True), (1, 255, False)],
                         ids=["f32", "f16r_colour", "f16"])
def test_checkpoint_resume_equals_uninterrupted(vh, oracle, wl, small_frames, tmp_path, storage, cap, color):
    from test_gpu_color import colour_frames

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    rgb = colour_frames(w, 20) if color else [None] * 20
    cfg = vh.config_from_workload(w, storage=storage, weight_cap=cap, color=int(color))
    print(f"checkpoint: storage {storage}, colour {color}", flush=True)
    whole = vh.Volume(cfg)
    for k in range(20):
        whole.integrate(depth[k], poses[k], rgb=rgb[k])
    for variant in ("resident", "host store"):
        part = vh.Volume(cfg)
        for k in range(10):
            part.integrate(depth[k], poses[k], rgb=rgb[k])
        if variant == "host store":  # half of the blocks saved from the host store
            n_out = part.stream_out(poses[9, :3, 3], 1.2)
            assert n_out > 100 and part.stats()["blocks_allocated"] > 100  # both parts substantial
        path = str(tmp_path / f"vol_{variant[0]}.vhv")
        part.save(path)
        part.close()
        vol = vh.Volume.load(path)
        st = vol.stats()
        assert st["host_blocks"] == 0 and st["frames"] == 10
        assert (vol.cfg.storage, vol.cfg.color, vol.cfg.pool_blocks) == (storage, int(color), cfg.pool_blocks)
        for k in range(10, 20):
            vol.integrate(depth[k], poses[k], rgb=rgb[k])
        a, b = _sorted_volume(oracle, vol, color), _sorted_volume(oracle, whole, color)
        assert np.array_equal(a[0], b[0]), variant
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8)), variant
        assert vol.stats()["frames"] == 20
        vol.close()
    whole.close()


def test_streaming_errors_leave_the_volume_unchanged(vh, oracle, wl, small_frames, tmp_path):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    probe = vh.Volume(vh.config_from_workload(w))
    for k in range(5):
        probe.integrate(depth[k], poses[k])
    n1 = probe.stats()["blocks_allocated"]
    probe.stream_out((0, 0, 0), 0.0)
    for k in range(30, 35):
        probe.integrate(depth[k], poses[k])
    n2 = probe.stats()["blocks_allocated"]
    probe.stream_in()
    n_union = probe.stats()["blocks_allocated"]
    probe.close()
    # the pool check counts only the keys that are not resident: a pool of exactly the union
    # takes the stored blocks (the others merge in place)
    exact = vh.Volume(vh.config_from_workload(w, pool_blocks=n_union))
    for k in range(5):
        exact.integrate(depth[k], poses[k])
    exact.stream_out((0, 0, 0), 0.0)
    for k in range(30, 35):
        exact.integrate(depth[k], poses[k])
    assert n1 + n2 > n_union and exact.stream_in() == n1
    assert exact.stats()["blocks_allocated"] == n_union
    exact.close()
    print(f"streaming errors: {n1} + {n2} blocks, pool {max(n1, n2) + 16}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, pool_blocks=max(n1, n2) + 16))
    for k in range(5):
        vol.integrate(depth[k], poses[k])
    assert vol.stream_out((0, 0, 0), 0.0) == n1
    for k in range(30, 35):
        vol.integrate(depth[k], poses[k])
    before = _sorted_volume(oracle, vol, False)
    with pytest.raises(vh.VhError) as e:
        vol.stream_in()  # n1 + n2 blocks do not fit
    assert e.value.code == vh.VH_ERR_POOL_FULL
    st = vol.stats()
    assert st["host_blocks"] == n1 and st["blocks_allocated"] == n2
    after = _sorted_volume(oracle, vol, False)
    assert all(np.array_equal(x, y) for x, y in zip(before, after))
    for centre, radius in (((math.nan, 0, 0), 1.0), ((0, 0, 0), -1.0), ((0, 0, 0), math.nan)):
        for fn in (vol.stream_out, vol.stream_in):
            with pytest.raises(vh.VhError) as e:
                fn(centre, radius)
            assert e.value.code == vh.VH_ERR_INVALID_ARG
    # a checkpoint cut short is refused
    path = tmp_path / "v.vhv"
    vol.save(str(path))
    data = path.read_bytes()
    path.write_bytes(data[: len(data) // 2])
    with pytest.raises(vh.VhError) as e:
        vh.Volume.load(str(path))
    assert e.value.code == vh.VH_ERR_IO
    vol.close()


def _same(got, want):
    keys, s, wt = got
    assert np.array_equal(keys, want[0]) and np.array_equal(wt, want[2])
    assert np.array_equal(s.view(np.uint32), want[1].view(np.uint32))


def _stream_cycle(vh, w, depth, poses, **over):
    """Frames 0-4, every block out, frames 5-9 (part of the keys allocated again)."""
    vol = vh.Volume(vh.config_from_workload(w, **over))
    for k in range(5):
        vol.integrate(depth[k], poses[k])
    n_a = vol.stream_out((0, 0, 0), 0.0)
    for k in range(5, 10):
        vol.integrate(depth[k], poses[k])
    return vol, n_a


def test_stream_out_and_save_keep_one_copy_per_key(vh, oracle, wl, small_frames, tmp_path):
    """A key that was streamed out and allocated again is in the pool and in the host store.
    Streaming it out again merges the stored copy into the resident block first (the same
    weighted mean as vh_stream_in), so nothing is lost; a checkpoint taken in that state holds
    each key once. Both end equal, bit for bit, to the volume that streamed everything back in."""
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    print("streaming: stream out and save with keys resident and stored", flush=True)
    ref, n_a = _stream_cycle(vh, w, depth, poses)
    n_b = ref.stats()["blocks_allocated"]
    ref.stream_in()
    n_union = ref.stats()["blocks_allocated"]
    n_both = ref.stats()["stream_merged_total"]
    assert n_both > 300 and n_union == n_a + n_b - n_both
    want = _sorted_volume(oracle, ref, False)

    # stream out again: one copy per key in the store, then everything back
    vol, _ = _stream_cycle(vh, w, depth, poses)
    assert vol.stream_out((0, 0, 0), 0.0) == n_b
    st = vol.stats()
    assert st["host_blocks"] == n_union and st["blocks_allocated"] == 0 and st["stream_merged_total"] == n_both
    path = tmp_path / "out.vhv"
    vol.save(str(path))
    assert vol.stream_in() == n_union
    _same(_sorted_volume(oracle, vol, False), want)
    vol.close()
    loaded = vh.Volume.load(str(path))
    _same(_sorted_volume(oracle, loaded, False), want)
    loaded.close()

    # save while keys are in both places: the checkpoint loads, with one copy per key
    vol, _ = _stream_cycle(vh, w, depth, poses)
    path2 = tmp_path / "both.vhv"
    vol.save(str(path2))
    assert vol.stats()["host_blocks"] == n_a - n_both
    vol.close()
    loaded = vh.Volume.load(str(path2))
    assert loaded.stats()["blocks_allocated"] == n_union
    _same(_sorted_volume(oracle, loaded, False), want)
    loaded.close()
    ref.close()
