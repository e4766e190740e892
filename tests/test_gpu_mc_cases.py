"""Marching cubes on chosen voxel values: every one of the 256 cube cases, on the GPU, against
the oracle.

A volume of 20 blocks (5 x 2 x 2) is written as a checkpoint (vh_save_volume's layout:
header, packed keys, FP32 voxel words) and loaded with vh_load_volume; the oracle gets the
same values through import_f32. The lattice is tiled by disjoint 2 x 2 x 2 cubes at even
coordinates, each given a case 0..255 (every case five times) with random magnitudes:
ordinary values, exact zeros and values at the 1e-5 snap threshold; about 3 % of the points
are unobserved (weight 0); the first cube of each case keeps ordinary values. Bar: the GPU's
triangle soup (cube records and vertex bits) and indexed mesh equal the oracle's bit for
bit, and the soup's cube records cover every non-trivial case.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEY_BIAS = 1 << 20


def pack_key(x, y, z):
    return ((x + KEY_BIAS) << 42) | ((y + KEY_BIAS) << 21) | (z + KEY_BIAS)


def write_checkpoint(vh, path, cfg, keys, tsdf, weight):
    """FP32 checkpoint: header (magic, ABI, 32-bit words per block, configuration, frame id,
    frame count, block count), packed keys, then per block 512 (tsdf, weight) pairs."""

    class Header(C.Structure):
        _fields_ = [("magic", C.c_char * 8), ("abi", C.c_int32), ("words_per_block", C.c_int32),
                    ("cfg", vh.Config), ("frame_id", C.c_uint32), ("reserved", C.c_uint32),
                    ("frames", C.c_uint64), ("n_blocks", C.c_int64)]

    h = Header()
    h.magic = b"VHVOLUME"
    h.abi = vh.ABI_VERSION
    h.words_per_block = 1024
    h.cfg = cfg
    h.frame_id = 1
    h.n_blocks = len(keys)
    packed = np.array([pack_key(*map(int, k)) for k in keys], np.uint64)
    words = np.stack([tsdf, weight], -1).astype(np.float32)
    with open(path, "wb") as f:
        f.write(bytes(h))
        f.write(packed.tobytes())
        f.write(words.tobytes())


def test_every_cube_case_matches_the_oracle(vh, oracle, wl, tmp_path):
    w = dict(wl.WORKLOADS["room_small_1cm"])
    cfg = vh.config_from_workload(w)
    rng = np.random.default_rng(256)
    nx, ny, nz = 40, 16, 16  # lattice points of the 5 x 2 x 2 blocks
    sign = np.zeros((nx, ny, nz), bool)  # True: value < 0 (corner inside)
    plain = np.zeros((nx, ny, nz), bool)  # the first cube of each case: observed, ordinary values
    corners = [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)]
    n = 0
    for i in range(0, nx, 2):
        for j in range(0, ny, 2):
            for k in range(0, nz, 2):
                for c, (dx, dy, dz) in enumerate(corners):
                    sign[i + dx, j + dy, k + dz] = bool(((n % 256) >> c) & 1)
                    plain[i + dx, j + dy, k + dz] = n < 256
                n += 1
    mag = rng.uniform(0.01, 1.0, sign.shape)
    special = np.where(plain, 1.0, rng.random(sign.shape))
    mag[special < 0.04] = 0.0  # exact zeros (on the positive side: value >= 0)
    mag[(special >= 0.04) & (special < 0.08)] = 1e-5
    mag[(special >= 0.08) & (special < 0.10)] = rng.uniform(0.5e-5, 2e-5, ((special >= 0.08) & (special < 0.10)).sum())
    val = np.where(sign, -np.maximum(mag, 1e-6), mag).astype(np.float32)
    wgt = np.where(~plain & (rng.random(sign.shape) < 0.03), 0.0, 1.0).astype(np.float32)
    keys, tsdf, weight = [], [], []
    for bx in range(5):
        for by in range(2):
            for bz in range(2):
                keys.append((bx, by, bz))
                tsdf.append(val[8 * bx:8 * bx + 8, 8 * by:8 * by + 8, 8 * bz:8 * bz + 8].reshape(512))
                weight.append(wgt[8 * bx:8 * bx + 8, 8 * by:8 * by + 8, 8 * bz:8 * bz + 8].reshape(512))
    keys, tsdf, weight = np.array(keys), np.array(tsdf), np.array(weight)
    path = str(tmp_path / "cases.vhv")
    write_checkpoint(vh, path, cfg, keys, tsdf, weight)
    print("mc cases: 20 blocks, 1,280 even cubes over the 256 cases", flush=True)
    vol = vh.Volume.load(path)
    assert vol.stats()["blocks_allocated"] == 20
    ref = oracle.Oracle(w, storage=0, weight_cap=0)
    ref.import_f32(keys, tsdf, weight)
    xyz, cubes = vol.extract_mesh()
    rxyz, rcubes = ref.mc()
    assert len(xyz) == len(rxyz) > 2000
    o, ro = np.lexsort(cubes.T[::-1]), np.lexsort(rcubes.T[::-1])
    assert np.array_equal(cubes[o], rcubes[ro]), "triangle topology differs"
    assert np.array_equal(xyz[o].view(np.uint32), rxyz[ro].view(np.uint32)), "vertex positions differ"
    assert set(np.unique(cubes[:, 3] >> 8).tolist()) == set(range(1, 255))
    v, f = vol.extract_mesh_indexed()
    rv, rf = ref.mesh_indexed()
    assert len(v) == len(rv) and len(f) == len(rf)
    rows = lambda a: a[np.lexsort(a.T[::-1])]
    assert np.array_equal(rows(v.view(np.uint32)), rows(rv.view(np.uint32)))
    assert np.array_equal(rows(v[f].reshape(-1, 9).view(np.uint32)), rows(rv[rf].reshape(-1, 9).view(np.uint32)))
    vol.close()
