"""GPU indexed mesh (vh_extract_mesh_indexed and its device variant) against the oracle.

Bar: the same vertices (float bits and colours) and the same faces (corner bits, in winding
order) as the oracle's indexed mesh, bit for bit, compared as sorted rows (vertex ids follow
pool indices, which concurrent allocation hands out in a different order on every run); a
closed sphere gives a closed genus-0 mesh; a mesh of several million faces equals the
oracle's too; the device variant reports counts, writes nothing when a buffer is too small,
and writes the arrays of the host variant.
"""
import numpy as np
import pytest

from conftest import sphere_views

pytestmark = pytest.mark.gpu


def rows(*cols):
    """Rows of per-item arrays as uint32 bit patterns (floats by their bits), sorted."""
    r = np.concatenate([np.ascontiguousarray(c).view(np.uint32) if c.dtype.kind == "f" else c.astype(np.uint32)
                        for c in cols], 1)
    return r[np.lexsort(r.T[::-1])]


def assert_same_indexed(got, want):
    (v, f, *c), (rv, rf, *rc) = got, want
    assert len(v) == len(rv) and len(f) == len(rf), (len(v), len(rv), len(f), len(rf))
    assert f.min() >= 0 and f.max() < len(v) and len(np.unique(f)) == len(v)
    assert np.array_equal(rows(v, *c), rows(rv, *rc)), "vertices differ"
    assert np.array_equal(rows(v[f].reshape(-1, 9)), rows(rv[rf].reshape(-1, 9))), "faces differ"
    if c:
        assert np.array_equal(rows(v[f].reshape(-1, 9), c[0][f].reshape(-1, 9)),
                              rows(rv[rf].reshape(-1, 9), rc[0][rf].reshape(-1, 9))), "vertex colours differ"


def edge_counts(f):
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    return np.unique(e, axis=0, return_counts=True)


def test_closed_sphere_is_closed_and_equals_the_oracle(vh, oracle, wl):
    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = sphere_views(w)
    print("indexed mesh: sphere, 6 views, FP32", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w, storage=0, weight_cap=0)
    for k in range(len(depth)):
        vol.integrate(depth[k], poses[k], scale=1e-4)
        ref.integrate(depth[k], poses[k], scale=1e-4)
    v, f = vol.extract_mesh_indexed()
    assert_same_indexed((v, f), ref.mesh_indexed())
    edges, counts = edge_counts(f)
    assert counts.min() == 2 and counts.max() == 2
    assert len(v) - len(edges) + len(f) == 2
    xyz, _ = vol.extract_mesh()
    assert 0 < len(f) - len(xyz) < 0.01 * len(f)  # the zero-area faces the soup drops
    vol.close()


@pytest.mark.parametrize("storage,cap", [(0, 0), (1, 255), (2, 255)], ids=["f32", "f16", "f16r"])
def test_room_colour_indexed_bit_exact(vh, oracle, wl, small_frames, storage, cap):
    from test_gpu_color import colour_frames

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    n = 12
    rgb = colour_frames(w, n)
    print(f"indexed mesh: room_small_1cm, {n} colour frames, storage {storage}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w, storage=storage, weight_cap=cap, color=1))
    ref = oracle.Oracle(w, storage=storage, weight_cap=cap, color=True)
    for k in range(n):
        vol.integrate(depth[k], poses[k], rgb=rgb[k])
        ref.integrate(depth[k], poses[k], rgb=rgb[k])
    got = vol.extract_mesh_indexed(with_rgb=True)
    assert len(got[1]) > 10000
    assert_same_indexed(got, ref.mesh_indexed(with_rgb=True))
    _, counts = edge_counts(got[1])
    assert counts.max() <= 2
    vol.close()


def test_large_mesh_and_device_variant(vh, oracle, wl):
    """room_medium_5mm, frames 0, 10, 20 (about 4.3 million faces): the host variant equals
    the oracle; the device variant reports the counts with zero capacities, writes nothing
    with one face or one vertex too few, and with exact capacities writes the host variant's
    arrays (same volume, same order)."""
    import torch

    w = wl.WORKLOADS["room_medium_5mm"]
    scene = vh.Scene(w["scene"], wl.SEED)
    depth, poses = scene.frames(0, 21)
    print("indexed mesh: room_medium_5mm, 3 frames", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    for k in (0, 10, 20):
        vol.integrate(depth[k], poses[k])
        ref.integrate(depth[k], poses[k])
    v, f = vol.extract_mesh_indexed()
    print(f"vertices {len(v)}, faces {len(f)}", flush=True)
    assert len(f) > (1 << 20) + (1 << 19)
    assert_same_indexed((v, f), ref.mesh_indexed())
    nv, nf, rc = vol.extract_mesh_indexed_device(None, 0, None, 0)
    assert rc == vh.VH_ERR_TRUNCATED and (nv, nf) == (len(v), len(f))
    d_v = torch.full((nv, 3), -1.0, dtype=torch.float32, device="cuda")
    d_f = torch.full((nf, 3), -1, dtype=torch.int32, device="cuda")
    _, _, rc = vol.extract_mesh_indexed_device(d_v, nv, d_f, nf - 1)
    assert rc == vh.VH_ERR_TRUNCATED
    _, _, rc = vol.extract_mesh_indexed_device(d_v, nv - 1, d_f, nf)
    assert rc == vh.VH_ERR_TRUNCATED
    assert (d_f == -1).all().item() and (d_v == -1).all().item()  # nothing written
    _, _, rc = vol.extract_mesh_indexed_device(d_v, nv, d_f, nf)
    assert rc == vh.VH_OK
    assert np.array_equal(d_v.cpu().numpy().view(np.uint32), v.view(np.uint32))
    assert np.array_equal(d_f.cpu().numpy(), f)
    vol.close()
