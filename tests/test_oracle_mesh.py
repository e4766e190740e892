"""Pins of the oracle's mesh measurement helpers (CPU only): the FP64 reference mesh
(marching cubes over the FP64 shadow) and the mesh-to-mesh distance behind the Hausdorff
figures.

Point-to-triangle distance is checked against an independent construction (the projection
onto the plane when it falls inside the triangle, else the nearest of the three clamped
segment projections); mesh distances against exact answers for identical and translated
meshes; the FP64 mesh of a fronto-parallel plane against its closed form.
"""
import numpy as np

from conftest import look_down_z, plane_depth


def _seg_dist(p, a, b):
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t * ab))


def _tri_dist_reference(p, a, b, c):
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n  # projection onto the plane
    # barycentric coordinates of q by least squares on the two edge vectors
    M = np.stack([b - a, c - a], 1)
    v, w = np.linalg.lstsq(M, q - a, rcond=None)[0]
    if v >= 0 and w >= 0 and v + w <= 1:
        return abs(np.dot(p - a, n))
    return min(_seg_dist(p, a, b), _seg_dist(p, b, c), _seg_dist(p, c, a))


def test_point_triangle_distance_matches_an_independent_construction(oracle):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        a, b, c = rng.normal(size=(3, 3))
        if np.linalg.norm(np.cross(b - a, c - a)) < 1e-3:
            continue
        p = rng.normal(size=3) * rng.choice([0.1, 1.0, 4.0])
        got = oracle.point_triangle_dist(p, a, b, c)
        want = _tri_dist_reference(p, a, b, c)
        assert abs(got - want) <= 1e-12 * max(1.0, want), (p, a, b, c, got, want)
    # points on the triangle (vertices, edge midpoints, centroid) are at distance 0
    a, b, c = np.array([0, 0, 0.0]), np.array([1, 0, 0.0]), np.array([0, 2, 0.0])
    for p in (a, b, c, (a + b) / 2, (b + c) / 2, (a + b + c) / 3):
        assert oracle.point_triangle_dist(p, a, b, c) < 1e-15


def _plane_oracle(oracle, wl, storage=0):
    w = dict(wl.WORKLOADS["room_small_1cm"])
    ref = oracle.Oracle(w, storage=storage, weight_cap=255 if storage else 0)
    ref.integrate(plane_depth(w, 1200), look_down_z(0.013, -0.021, 0.0))
    return w, ref


def test_fp64_mesh_of_a_plane_is_the_plane(oracle, wl):
    """A fronto-parallel plane at z = 1.2 m: the FP64 shadow is exactly (d - z)/mu, so every
    vertex of the FP64 mesh lies on z = 1.2 m up to the FP32 rounding of the output; the FP32
    twin's mesh lies within a few 1e-6 m of it."""
    w, ref = _plane_oracle(oracle, wl)
    xyz64, c64 = ref.mc64()
    xyz, c = ref.mc()
    assert len(xyz64) > 10000
    z64 = xyz64.reshape(-1, 3)[:, 2].astype(np.float64)
    assert np.abs(z64 - 1.2).max() <= 1.2 * 2.0 ** -23
    z32 = xyz.reshape(-1, 3)[:, 2].astype(np.float64)
    assert np.abs(z32 - 1.2).max() < 5e-6
    # the same cubes carry the surface in both meshes (no sign flips on a clean plane)
    assert np.array_equal(np.unique(c64[:, :3], axis=0), np.unique(c[:, :3], axis=0))


def test_mesh_distance_identity_and_translation(oracle, wl):
    w, ref = _plane_oracle(oracle, wl)
    xyz, cubes = ref.mc64()
    d = oracle.mesh_distance(xyz, cubes, xyz, cubes)
    assert d["max"] < 1e-12 and d["far"] == 0 and d["samples"] == 4 * len(xyz)  # centroids round
    # the plane moved along its normal by 0.3 voxel: every sample is 0.3 voxel away
    delta = 0.3 * w["vox_size"]
    moved = xyz.copy()
    moved[:, 2::3] += np.float32(delta)
    h = oracle.hausdorff(xyz, cubes, moved, cubes)
    assert abs(h["hausdorff"] - delta) < 1e-6
    assert abs(h["a_to_b"]["mean"] - delta) < 1e-6 and h["a_to_b"]["far"] == 0
    # a mesh compared with its own half: the missing half is "far", the rest at distance 0
    half = len(xyz) // 2
    d = oracle.mesh_distance(xyz[:half], cubes[:half], xyz, cubes)
    assert d["max"] < 1e-12
    d = oracle.mesh_distance(xyz, cubes, xyz[:half], cubes[:half])
    assert d["far"] > 0
