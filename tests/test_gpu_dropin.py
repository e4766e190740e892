"""The C++ drop-in (ark::GpuTsdfGenerator over the C ABI) on the GPU: the example caller
integrates frames and writes the reference's ASCII PLY layout with shared vertices; both PLY
writers carry the library's indexed mesh."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def read_ascii_ply(path, with_rgb=False):
    with open(path) as f:
        lines = f.read().splitlines()
    end = lines.index("end_header")
    head = lines[:end]
    nv = int([l for l in head if l.startswith("element vertex")][0].split()[-1])
    nf = int([l for l in head if l.startswith("element face")][0].split()[-1])
    rows = [l.split() for l in lines[end + 1:end + 1 + nv]]
    v = np.array([list(map(float, r[:3])) for r in rows])
    f = np.array([list(map(int, l.split()[1:])) for l in lines[end + 1 + nv:end + 1 + nv + nf]])
    if with_rgb:
        return v, f, np.array([list(map(int, r[3:6])) for r in rows], np.int64)
    return v, f


def _rows(*cols):
    """Rows of the given per-item arrays as uint32 bit patterns (floats by their bits), sorted."""
    r = np.concatenate([np.ascontiguousarray(c).astype(np.float32).view(np.uint32) if c.dtype.kind == "f"
                        else c.astype(np.uint32) for c in cols], 1)
    return r[np.lexsort(r.T[::-1])]


def test_load_frames_example_writes_indexed_ply(vh, wl, small_frames, tmp_path):
    exe = os.path.join(ROOT, "voxel-hashing-sdf_amd", "load_frames")
    out = tmp_path / "room.ply"
    print("drop-in: load_frames, 10 frames with colour", flush=True)
    r = subprocess.run([exe, str(out), "10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v, f, vrgb = read_ascii_ply(out, with_rgb=True)
    assert len(v) > 1000 and len(f) > 1000
    assert f.min() >= 0 and f.max() < len(v)
    # the same frames through the Python binding: the PLY holds the library's indexed mesh
    # (vertex order follows pool indices, which differ between runs: compared as sorted rows
    # of vertex bits and colours, and of face corner bits)
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    scene = vh.Scene(w["scene"], wl.SEED)
    vol = vh.Volume(vh.config_from_workload(w, log2_buckets=20, pool_blocks=1 << 19, color=1, carve=1))
    for k in range(10):
        vol.integrate(depth[k].astype(np.float32) * np.float32(0.001), poses[k], rgb=scene.rgb(k, poses[k]))
    iv, fi, irgb = vol.extract_mesh_indexed(with_rgb=True)
    assert len(fi) == len(f) and len(iv) == len(v)
    assert np.array_equal(_rows(v, vrgb), _rows(iv, irgb))
    assert np.array_equal(_rows(v.astype(np.float32)[f].reshape(-1, 9)), _rows(iv[fi].reshape(-1, 9)))
    assert len({tuple(c) for c in vrgb}) > 100  # textured, not constant
    # every vertex is used, and each edge lies in two faces at most (a manifold mesh)
    assert len(np.unique(f)) == len(v)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert counts.max() <= 2
    # the soup's triangles are the faces of positive area
    xyz, _ = vol.extract_mesh()
    assert len(fi) - len(xyz) < 0.01 * len(xyz)
    # binary PLY of the library: the same volume, so the same arrays
    bpath = tmp_path / "room_bin.ply"
    vol.save_ply(str(bpath))
    raw = bpath.read_bytes()
    hdr_end = raw.index(b"end_header\n") + len(b"end_header\n")
    hdr = raw[:hdr_end].decode()
    assert "property uchar red\nproperty uchar green\nproperty uchar blue\n" in hdr
    assert f"element vertex {len(iv)}\n" in hdr and f"element face {len(fi)}\n" in hdr
    rec = np.frombuffer(raw[hdr_end:hdr_end + len(iv) * 15], np.dtype([("p", "<f4", 3), ("c", "u1", 3)]))
    assert np.array_equal(rec["p"].view(np.uint32), iv.view(np.uint32))
    assert np.array_equal(rec["c"], irgb)
    frec = np.frombuffer(raw[hdr_end + len(iv) * 15:], np.dtype([("n", "u1"), ("i", "<i4", 3)]))
    assert len(frec) == len(fi) and (frec["n"] == 3).all() and np.array_equal(frec["i"], fi)
    vol.close()


def test_dataset_mode_equals_synthetic_mode(vh, small_frames, tmp_path):
    """The frames of the synthetic path written as a reference-style dataset (depth/<id>.png,
    tcw/<id>.txt, settings file) and integrated by load_frames --dataset give the same mesh as
    load_frames integrating them directly: the same triangles with the same vertex bits (the
    order in the file follows pool indices, which concurrent allocation hands out in a
    different order on every run)."""
    exe = os.path.join(ROOT, "voxel-hashing-sdf_amd", "load_frames")
    depth, poses = small_frames
    n = 8
    ds = tmp_path / "ds"
    (ds / "depth").mkdir(parents=True)
    (ds / "tcw").mkdir()
    for k in range(n):
        vh.png16_save(str(ds / "depth" / f"{k}.png"), depth[k])
        # 9 significant digits give each float32 entry back exactly
        (ds / "tcw" / f"{k}.txt").write_text("\n".join(" ".join(f"{x:.9g}" for x in row) for row in poses[k]) + "\n")
    settings = tmp_path / "cam.yaml"
    settings.write_text("%YAML:1.0\nCamera.fx: 525\nCamera.fy: 525\nCamera.cx: 319.5\nCamera.cy: 239.5\n"
                        "Camera.width: 640\nCamera.height: 480\nDepthMapFactor: 1000\nMaxDepth: 5.0\n"
                        "Voxel.Origin.x: -1.5\nVoxel.Origin.y: -1.5\nVoxel.Origin.z: 0.5\nVoxel.Size: 0.01\n"
                        "Voxel.TruncMargin: 0.04\nVoxel.Dim.x: 500\nVoxel.Dim.y: 500\nVoxel.Dim.z: 500\n")
    print(f"drop-in: load_frames --dataset, {n} frames", flush=True)
    a, b = tmp_path / "direct.ply", tmp_path / "dataset.ply"
    r = subprocess.run([exe, str(a), str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, str(b), "--dataset", str(ds), str(settings)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"frames {n} " in r.stdout

    def triangles(path):
        """One row of 9 float bit patterns per face (corners in the face's order), rows sorted."""
        v, f = read_ascii_ply(path)
        rows = v.astype(np.float32).view(np.uint32)[f].reshape(-1, 9)
        return rows[np.lexsort(rows.T[::-1])]

    ta, tb = triangles(a), triangles(b)
    assert ta.shape == tb.shape and len(ta) > 1000
    assert np.array_equal(ta, tb)


def test_streaming_generator_writes_the_same_mesh(tmp_path):
    """ark::GpuTsdfGenerator::setStreaming (stream-out every frame, no margin) over the 40
    small-room frames with a 1.2 m depth range: blocks leave the GPU and come back, and the
    PLY holds the same triangles, with the same vertex and colour bits, as without streaming."""
    exe = os.path.join(ROOT, "voxel-hashing-sdf_amd", "load_frames")
    a, b = tmp_path / "plain.ply", tmp_path / "streamed.ply"
    print("drop-in: 40 frames, max depth 1.2 m, with and without streaming", flush=True)
    r = subprocess.run([exe, str(a), "40", "--max-depth", "1.2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, str(b), "40", "--max-depth", "1.2", "--stream"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    streamed = int(r.stdout.split("streamed_out")[1].split()[0])
    print(r.stdout.strip(), flush=True)
    assert streamed > 0

    def triangles(path):
        v, f, c = read_ascii_ply(path, with_rgb=True)
        rows = np.concatenate([v.astype(np.float32).view(np.uint32)[f].reshape(-1, 9),
                               c[f].reshape(-1, 9).astype(np.uint32)], 1)
        return rows[np.lexsort(rows.T[::-1])]

    ta, tb = triangles(a), triangles(b)
    assert len(ta) > 1000 and np.array_equal(ta, tb)


def test_every_dropin_member_on_the_gpu(vh, tmp_path):
    """The caller of every public member (tests/host_checks/dropin_api.cpp): a plane 1 m in
    front of the camera, three frames. Its PLY holds the library's indexed mesh; SaveTSDF's
    dense grid holds the exported voxel values at their lattice positions (1 and colour 0
    elsewhere); find_vertex finds every vertex inside the hash grid."""
    from test_abi_host import build_dropin_api

    exe = build_dropin_api(tmp_path)
    print("drop-in: every public member, plane at 1 m", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    tok = r.stdout.split()
    assert tok[0] == "ok", r.stdout
    nv, nf, found = tok[2], tok[4], tok[6]
    v, f = read_ascii_ply(tmp_path / "plane.ply")
    assert len(v) == int(nv) > 1000 and len(f) == int(nf)
    inside = np.all((v >= 0) & (v < 2.56), 1).sum()
    assert int(found) == inside > 1000
    # the same frames through the binding
    cfg = dict(width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, max_depth=3.0, vox_size=0.01,
               trunc=0.04, log2_buckets=20, pool_blocks=1 << 19, color=1, carve=1)
    c = vh.make_config(**cfg)
    c.origin[:] = (-1.5, -1.5, 0.5)
    c.vox_dim[:] = (256, 256, 256)
    vol = vh.Volume(c)
    depth = np.ones((480, 640), np.float32)
    rgb = (np.arange(480 * 640 * 3) * 7 % 251).astype(np.uint8).reshape(480, 640, 3)
    T = np.eye(4, dtype=np.float32)
    T[:2, 3] = 1.28
    for _ in range(3):
        vol.integrate(depth, T, rgb=rgb)
    keys, tsdf, wt = vol.export_blocks()
    cols = vol.export_block_colors()
    vol.close()
    raw = open(tmp_path / "plane.tsdf", "rb").read()
    head = np.frombuffer(raw[:32], np.float32)
    assert list(head) == [256, 256, 256, np.float32(-1.5), np.float32(-1.5), np.float32(0.5), np.float32(0.01),
                          np.float32(0.04)]
    n = 256 ** 3
    grid = np.frombuffer(raw[32:32 + 4 * n], np.float32).reshape(256, 256, 256)  # z, y, x
    cgrid = np.frombuffer(raw[32 + 4 * n:], np.uint8).reshape(256, 256, 256, 3)
    i = np.arange(512)
    lat = keys[:, None, :] * 8 + np.stack([i >> 6, (i >> 3) & 7, i & 7], 1)[None] - np.array([-150, -150, 50])
    ok = np.all((lat >= 0) & (lat < 256), 2) & (wt > 0)
    x, y, z = lat[ok].T
    assert ok.sum() > 5000
    assert np.array_equal(grid[z, y, x].view(np.uint32), tsdf[ok].view(np.uint32))
    assert np.array_equal(cgrid[z, y, x], cols[ok])
    mask = np.zeros((256, 256, 256), bool)
    mask[z, y, x] = True
    assert np.all(grid[~mask] == 1.0) and not cgrid[~mask].any()
