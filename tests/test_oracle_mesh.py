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


def _edges_of(faces):
    e = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), 1)
    return np.unique(e, axis=0, return_counts=True)


_CORNER = np.array([[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1], [1, 0, 1]])
_EDGE = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def _sphere_oracle(oracle, wl):
    from conftest import sphere_views

    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = sphere_views(w)
    ref = oracle.Oracle(w, storage=0, weight_cap=0)
    for k in range(len(depth)):
        ref.integrate(depth[k], poses[k], scale=1e-4)
    return w, ref


def _crossed_edges(ref):
    """Lattice edges whose two end points are observed and on opposite sides (< 0 vs >= 0),
    as (lower point, axis), from the exported volume."""
    keys, s, wt, _, _ = ref.export()
    i = np.arange(512)
    local = np.stack([i >> 6, (i >> 3) & 7, i & 7], 1)
    pts = (keys[:, None, :] * 8 + local[None]).reshape(-1, 3)
    obs = wt.reshape(-1) > 0
    pts, neg = pts[obs], s.reshape(-1)[obs] < 0
    at = {tuple(p): n for p, n in zip(pts.tolist(), neg.tolist())}
    out = set()
    for p, n in at.items():
        for ax in range(3):
            q = list(p)
            q[ax] += 1
            m = at.get(tuple(q))
            if m is not None and m != n:
                out.add((*p, ax))
    return out


def test_indexed_mesh_of_a_closed_sphere(oracle, wl):
    """Six analytic views of a sphere: the oracle's indexed mesh is a closed 2-manifold of
    genus 0 (every edge in exactly two faces, V - E + F = 2); its vertices are exactly the
    crossed lattice edges of the volume (end points observed, opposite signs) and the keys
    its faces name (counted here from the face cube records and the triangle table); its
    faces are the soup's triangles plus zero-area ones, with each soup corner where the soup
    put it (equal, or a rounding apart where the soup interpolated from the upper end)."""
    w, ref = _sphere_oracle(oracle, wl)
    xyz, cubes = ref.mc()
    v, f, fc = ref.mesh_indexed(with_cubes=True)
    assert len(f) >= len(xyz) > 5000
    edges, counts = _edges_of(f)
    assert counts.min() == 2 and counts.max() == 2, np.bincount(counts)
    assert len(v) - len(edges) + len(f) == 2
    keys = []
    for c in fc:
        s = oracle.tri_case(int(c[3]) >> 8)
        t = int(c[3]) & 255
        for j in range(3):
            a, b = _EDGE[int(s[3 * t + j], 16)]
            lo = c[:3] + np.minimum(_CORNER[a], _CORNER[b])
            keys.append((*lo.tolist(), int(np.argmax(_CORNER[a] != _CORNER[b]))))
    assert len(set(keys)) == len(v) == len(_crossed_edges(ref))
    assert set(keys) == _crossed_edges(ref)
    # faces naming the same key share the vertex id, distinct keys distinct ids
    ids = {}
    for k, i in zip(keys, f.reshape(-1).tolist()):
        assert ids.setdefault(k, i) == i
    assert len(set(ids.values())) == len(ids)
    # soup triangles are the faces with the same cube record, in order
    rec = [tuple(c) for c in fc.tolist()]
    pos = {r: i for i, r in enumerate(rec)}
    sel = np.array([pos[tuple(c)] for c in cubes.tolist()])
    assert np.all(np.diff(sel) > 0)
    soup = xyz.reshape(-1, 3).astype(np.float64)
    idx = v[f[sel].reshape(-1)].astype(np.float64)
    diff = np.abs(idx - soup)
    assert np.all(diff <= 2.0 ** -22 * np.maximum(np.abs(soup), w["vox_size"])), diff.max()
    assert (diff == 0).mean() > 0.5
    # the faces the soup leaves out have two coincident corners (up to a rounding)
    extra = np.setdiff1d(np.arange(len(f)), sel)
    assert 0 < len(extra) < 0.01 * len(f)
    tri = v[f[extra]].astype(np.float64)
    short = np.stack([np.abs(tri[:, a] - tri[:, b]).max(1) for a, b in ((0, 1), (1, 2), (2, 0))], 1).min(1)
    assert short.max() <= 2.0 ** -20, short.max()
