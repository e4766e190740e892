"""Pins the oracle's CPU scatter of the voxel gradient (vho_scatter_grad), the reference for the
GPU backward's d/d(voxel SDF): the sample is linear in the voxel values, so the derivative of
sum_i up_i r_i with respect to one voxel value equals a finite difference of that value exactly
(up to FP64 rounding); the eight weights of a sample sum to one; a point on a lattice vertex
puts its whole weight on that voxel.
"""
import numpy as np

VOX = 0.01


def _volume(oracle, wl, small_frames, n=10):
    w = dict(wl.WORKLOADS["room_small_1cm"])
    depth, poses = small_frames
    ref = oracle.Oracle(w)
    for k in range(n):
        ref.integrate(depth[k], poses[k])
    return w, ref


def _points(ref, n, seed):
    """Random points inside observed cells near the surface."""
    keys, s, wt, _, _ = ref.export()
    rng = np.random.default_rng(seed)
    b, i = np.nonzero((wt > 0) & (np.abs(s) < 0.5))
    pick = rng.choice(len(b), n)
    lat = keys[b[pick]] * 8 + np.stack([i[pick] >> 6, (i[pick] >> 3) & 7, i[pick] & 7], 1)
    f32 = lambda x: float(np.float32(x))
    return (lat + rng.uniform(0, 1, (n, 3))) * f32(VOX)


def test_scatter_weights_sum_to_the_upstream_and_lattice_points(oracle, wl, small_frames):
    _, ref = _volume(oracle, wl, small_frames)
    pts = _points(ref, 20000, 1)
    up = np.random.default_rng(2).normal(size=len(pts)).astype(np.float32)
    keys, g, used = ref.scatter_grad(pts, up)
    _, _, valid = ref.query64(pts)
    assert used == valid.sum() > 10000
    assert abs(g.sum() - up[valid].astype(np.float64).sum()) < 1e-9 * np.abs(up).sum()
    # a point on a lattice vertex: weight 1 on that voxel, 0 on the others
    k0, s0, w0, _, _ = ref.export()
    b, i = np.argwhere((w0 > 0) & (np.abs(s0) < 0.5))[0]
    lat = k0[b] * 8 + np.array([i >> 6, (i >> 3) & 7, i & 7])
    p = (lat.astype(np.float64) + 1e-9) * float(np.float32(VOX))
    _, _, ok = ref.query64(p[None])
    if ok[0]:
        _, g1, u1 = ref.scatter_grad(p[None], np.array([1.0], np.float32))
        assert u1 == 1 and abs(g1[b, i] - 1.0) < 1e-6 and abs(g1.sum() - 1.0) < 1e-12


def test_scatter_equals_finite_differences_of_single_voxel_values(oracle, wl, small_frames):
    _, ref = _volume(oracle, wl, small_frames)
    pts = _points(ref, 20000, 3)
    up = np.random.default_rng(4).normal(size=len(pts)).astype(np.float32)
    keys, g, _ = ref.scatter_grad(pts, up)
    k0, s0, w0, _, _ = ref.export()
    assert np.array_equal(keys, k0)
    wl_ = dict(wl.WORKLOADS["room_small_1cm"])
    rng = np.random.default_rng(5)
    cand = np.argwhere(np.abs(g) > 1e-3)
    for b, i in cand[rng.choice(len(cand), 6, replace=False)]:
        r = []
        for sign in (1, -1):
            s = s0.copy()
            s[b, i] = np.float32(s0[b, i] + sign * np.float32(1e-2))
            o = oracle.Oracle(wl_, storage=0)
            o.import_f32(k0, s, w0)
            val, _, ok = o.query64(pts)
            r.append((np.where(ok, val, 0.0), s[b, i]))
        (rp, sp), (rm, sm) = r
        fd = (up.astype(np.float64) * (rp - rm)).sum() / (float(sp) - float(sm))
        assert abs(fd - g[b, i]) < 1e-9 * max(1.0, abs(g[b, i])), (b, i, fd, g[b, i])


def test_twin_scatter_differs_from_fp64_by_weight_rounding(oracle, wl, small_frames):
    """The twin scatter (FP32 weights and products, as the kernels form them) against the FP64
    one: per voxel within 4e-7 of sum_i |up_i| over its contributions (FP32 rounding of values
    at most 1, a few roundings each)."""
    _, ref = _volume(oracle, wl, small_frames)
    pts = _points(ref, 20000, 6)
    up = np.random.default_rng(7).normal(size=len(pts)).astype(np.float32)
    _, g64, u1 = ref.scatter_grad(pts, up)
    _, g32, u2 = ref.scatter_grad(pts, up, twin=True)
    _, gabs, _ = ref.scatter_grad(pts, np.abs(up))
    assert u1 == u2
    assert np.abs(g64 - g32).max() > 0
    # sum_i |up_i| over the points touching a voxel is at least gabs (weights <= 1) and at
    # most 8x larger for points near a corner; bound by 4e-7 * 8 gabs plus the smallest scale
    assert np.all(np.abs(g64 - g32) <= 4e-7 * (gabs + np.abs(up).max()))
