"""The C++ drop-in (ark::GpuTsdfGenerator over the C ABI) on the GPU: the example caller
integrates frames and writes the reference's ASCII PLY layout; welding and the binary PLY
writer agree with the library's triangle soup."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def read_ascii_ply(path):
    with open(path) as f:
        lines = f.read().splitlines()
    end = lines.index("end_header")
    head = lines[:end]
    nv = int([l for l in head if l.startswith("element vertex")][0].split()[-1])
    nf = int([l for l in head if l.startswith("element face")][0].split()[-1])
    v = np.array([list(map(float, l.split()[:3])) for l in lines[end + 1:end + 1 + nv]])
    f = np.array([list(map(int, l.split()[1:])) for l in lines[end + 1 + nv:end + 1 + nv + nf]])
    return v, f


def test_load_frames_example_writes_welded_ply(vh, wl, small_frames, tmp_path):
    exe = os.path.join(ROOT, "voxel-hashing-sdf_amd", "load_frames")
    out = tmp_path / "room.ply"
    print("drop-in: load_frames, 10 frames", flush=True)
    r = subprocess.run([exe, str(out), "10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v, f = read_ascii_ply(out)
    assert len(v) > 1000 and len(f) > 1000
    assert f.min() >= 0 and f.max() < len(v)
    # the same frames through the Python binding: as many faces as triangles, and the welded
    # vertices are the distinct triangle corners
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w, log2_buckets=20, pool_blocks=1 << 19))
    for k in range(10):
        vol.integrate(depth[k].astype(np.float32) * np.float32(0.001), poses[k])
    xyz, _ = vol.extract_mesh()
    assert len(xyz) == len(f)
    corners = np.unique(xyz.reshape(-1, 3), axis=0)
    assert len(corners) == len(v)
    # each edge of a closed patch is shared by two faces at most (welded, manifold soup)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert counts.max() <= 2
    # binary PLY of the library
    bpath = tmp_path / "room_bin.ply"
    vol.save_ply(str(bpath))
    raw = bpath.read_bytes()
    hdr_end = raw.index(b"end_header\n") + len(b"end_header\n")
    nv = len(xyz) * 3
    verts = np.frombuffer(raw[hdr_end:hdr_end + nv * 12], np.float32).reshape(-1, 9)
    assert np.array_equal(np.sort(verts.view(np.uint32), 0), np.sort(xyz.view(np.uint32), 0))
    vol.close()
