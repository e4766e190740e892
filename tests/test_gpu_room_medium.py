"""GPU parity at the medium room's full size (room_medium_5mm: 500 frames of an 8 x 6 x 3 m room
with 12 primitives, 5 mm voxels, 2 cm truncation, 2^21 buckets, 1.5M-block pool), as the
benchmark runs it: the frames bound as a device sequence and integrated in batches of 8.

- FP32 with marching cubes every 100 frames: the volume bit-identical to the oracle's FP32
  twin after 500 frames, the triangle count of every periodic extraction equal to the
  oracle's at the same frame, the final mesh identical.
- FP16 (binary16 tsdf + uint16 weight, cap 255): bit-identical to the oracle's FP16 twin; its
  error against the oracle's FP64 shadow stays within the floor of an 11-bit significand
  (measured max 9.3e-3, RMS 7.9e-4; bounds 1.5e-2 and 1.5e-3: the weight-255 mean's increments
  fall below half an ulp, tests/test_storage_precision.py), so it does not meet 1e-3; its
  mesh lies within 0.1 mm of the FP64 mesh at the 99.9th percentile. (The largest
  distance, 1.7 cm, comes from noise specks whose tiny closed surfaces change sign between
  precisions; FP32 has them too, 4.8 mm.)
- F16R (binary16 + int8 residual + uint8 weight): bit-identical to its twin, every observed
  voxel within 1e-3 of the FP64 shadow (the 4-byte format meant for that bound).
- Differentiation on the FP16 volume: all valid pixels of the held-out frame 517 (about
  300k points) under a seeded camera-frame pose error (3 cm, 2 degrees per axis): per-point
  Jacobians against the FP64 analytic ones and central differences (h = 1e-4) within 1e-5;
  the voxel gradient of sum_i up_i r_i against the oracle's CPU scatter (itself pinned to
  finite differences of single voxel values, tests/test_oracle_grad.py); 10 Levenberg-
  Marquardt iterations within 1e-6 of the CPU FP64 restatement, from the full error and from a
  quarter of it, converging without drift after the fourth evaluation.
"""
import numpy as np
import pytest

from gn_reference import lm_reference, pose_error, se3_exp, world_points

pytestmark = pytest.mark.gpu

NF = 500


@pytest.fixture(scope="module")
def medium(vh, wl):
    w = dict(wl.WORKLOADS["room_medium_5mm"])
    scene = vh.Scene(w["scene"], wl.SEED)
    depth, poses = scene.frames(0, NF)
    return w, scene, depth, poses


def _sequence_volume(vh, w, depth, poses, mc_every=0, **over):
    import torch

    vol = vh.Volume(vh.config_from_workload(w, **over))
    d_depth = torch.from_numpy(np.ascontiguousarray(depth).view(np.int16)).cuda()
    d_pose = torch.from_numpy(np.ascontiguousarray(poses, np.float32).reshape(len(poses), 16)).cuda()
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, len(depth))
    counts = []
    if mc_every:
        cap = 16 << 20
        d_xyz = torch.empty(cap * 9, dtype=torch.float32, device="cuda")
        for f in range(0, len(depth), mc_every):
            vol.integrate_next(min(mc_every, len(depth) - f))
            counts.append(vol.extract_mesh_device(d_xyz, None, cap))
        del d_xyz
    else:
        vol.integrate_next(len(depth))
    vol.sync()
    return vol, counts, (d_depth, d_pose)


def _compare(oracle, vol, ref):
    keys, s, wt = vol.export_blocks()
    keys, s, wt = oracle.sort_blocks(keys, s, wt)
    rk, rs, rw, r64, _ = ref.export()
    assert np.array_equal(keys, rk), "allocated block sets differ"
    assert np.array_equal(wt, rw), "weights differ"
    bad = s.view(np.uint32) != rs.view(np.uint32)
    assert not bad.any(), f"{bad.sum()} tsdf values differ"
    m = rw > 0
    e = np.abs(s[m].astype(np.float64) - r64[m])
    return float(e.max()), float(np.sqrt((e ** 2).mean()))


def test_fp32_500_frames_with_periodic_marching_cubes(vh, oracle, medium):
    w, _, depth, poses = medium
    print(f"room_medium_5mm: {NF} frames FP32, batches of 8, MC every 100 frames", flush=True)
    vol, counts, _ = _sequence_volume(vh, w, depth, poses, mc_every=100)
    ref = oracle.Oracle(w, storage=0)
    want = []
    for k in range(NF):
        ref.integrate(depth[k], poses[k])
        if k % 100 == 99:
            want.append(len(ref.mc()[0]))
    print(f"triangles every 100 frames: device {counts}, oracle {want}", flush=True)
    assert counts == want
    emax, erms = _compare(oracle, vol, ref)
    print(f"FP32 vs FP64 shadow: max {emax:.2e} rms {erms:.2e}")
    xyz, cubes = vol.extract_mesh()
    rxyz, rcubes = ref.mc()
    assert len(xyz) == len(rxyz) == counts[-1]
    o = np.lexsort(cubes.T[::-1])
    ro = np.lexsort(rcubes.T[::-1])
    assert np.array_equal(cubes[o], rcubes[ro])
    assert np.array_equal(xyz[o].view(np.uint32), rxyz[ro].view(np.uint32))
    vol.close()


@pytest.fixture(scope="module")
def fp16_volume(vh, oracle, medium):
    w, _, depth, poses = medium
    print(f"room_medium_5mm: {NF} frames FP16 (binary16 + uint16 weight, cap 255)", flush=True)
    vol, _, keep = _sequence_volume(vh, w, depth, poses, storage=1, weight_cap=255)
    ref = oracle.Oracle(w, storage=1, weight_cap=255)
    for k in range(NF):
        ref.integrate(depth[k], poses[k])
    yield w, vol, ref
    vol.close()


def test_fp16_500_frames_bit_exact_and_its_error_floor(oracle, fp16_volume):
    w, vol, ref = fp16_volume
    emax, erms = _compare(oracle, vol, ref)
    xyz, cubes = vol.extract_mesh()
    h = oracle.hausdorff(xyz, cubes, *ref.mc64())
    print(f"FP16 vs FP64 shadow: max {emax:.2e} rms {erms:.2e}; Hausdorff to the FP64 mesh "
          f"{1e3 * h['hausdorff']:.3f} mm (q0.99999 {1e3 * h['q0.99999']:.3f} mm)")
    assert emax < 1.5e-2 and erms < 1.5e-3
    assert emax > 1e-3  # the format does not meet the metric's 1e-3 (F16R does)
    assert h["q0.999"] < 1e-4 and h["hausdorff"] < w["trunc"]


def test_f16r_500_frames_meets_1e3(vh, oracle, medium):
    w, _, depth, poses = medium
    print(f"room_medium_5mm: {NF} frames F16R", flush=True)
    vol, _, _ = _sequence_volume(vh, w, depth, poses, storage=2, weight_cap=255)
    ref = oracle.Oracle(w, storage=2, weight_cap=255)
    for k in range(NF):
        ref.integrate(depth[k], poses[k])
    emax, erms = _compare(oracle, vol, ref)
    print(f"F16R vs FP64 shadow: max {emax:.2e} rms {erms:.2e}")
    assert emax <= 1e-3
    vol.close()


def _held_out(vh, wl, w):
    scene = vh.Scene(w["scene"], wl.SEED)
    d, T = scene.frames(NF + 17, 1)
    f32 = np.float32
    dep = d[0].astype(f32) * f32(0.001)
    v, u = np.mgrid[0:dep.shape[0], 0:dep.shape[1]]
    ok = (dep > 0) & (dep <= f32(w["max_depth"]))
    x = ((u.astype(f32) - f32(w["cx"])) * dep) / f32(w["fx"])
    y = ((v.astype(f32) - f32(w["cy"])) * dep) / f32(w["fy"])
    pts = np.stack([x[ok], y[ok], dep[ok]], 1).astype(f32)
    rng = np.random.default_rng(wl.SEED + 2)
    xi = np.concatenate([rng.normal(0, 0.03, 3), rng.normal(0, np.deg2rad(2.0), 3)])
    return pts, T[0].astype(np.float64), xi


def test_backward_on_the_fp16_volume(vh, oracle, wl, fp16_volume):
    import torch

    w, vol, ref = fp16_volume
    pts, T_true, xi = _held_out(vh, wl, w)
    T = (T_true @ se3_exp(xi)).astype(np.float32)
    n = len(pts)
    print(f"backward: {n} points of held-out frame {NF + 17} at the perturbed pose", flush=True)
    assert n > 250_000
    rng = np.random.default_rng(11)
    up = rng.normal(size=n).astype(np.float32)
    d_pts = torch.from_numpy(pts).cuda()
    d_up = torch.from_numpy(up).cuda()
    d_r = torch.zeros(n, dtype=torch.float32, device="cuda")
    d_valid = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_J = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    d_pg = torch.zeros(6, dtype=torch.float64, device="cuda")
    d_vg = torch.zeros(w["pool_blocks"] * 512, dtype=torch.float32, device="cuda")
    vol.backward_device(d_pts, n, T, d_up, d_r, d_valid, d_J, d_pg, d_vg)
    torch.cuda.synchronize()
    valid = d_valid.cpu().numpy().astype(bool)
    J = d_J.cpu().numpy()
    q = world_points(T, pts)
    s64, g64, v64 = ref.query64(q)
    assert np.array_equal(valid, v64) and valid.sum() > 100_000
    J64 = np.concatenate([g64, np.cross(q, g64)], 1)
    rel = np.linalg.norm(J[valid] - J64[valid]) / np.linalg.norm(J64[valid])
    print(f"J vs FP64 analytic: {rel:.2e}")
    assert rel < 1e-5
    # central differences (h = 1e-4) of the FP64 sample, points whose stencil stays in a cell
    h, vox = 1e-4, float(np.float32(w["vox_size"]))
    q64 = q[valid][:30000]
    cell = np.floor(q64 / vox)
    ok = np.ones(len(q64), bool)
    fd = np.zeros((len(q64), 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Ep, Em = se3_exp(e), se3_exp(-e)
        qp = q64 @ Ep[:3, :3].T + Ep[:3, 3]
        qm = q64 @ Em[:3, :3].T + Em[:3, 3]
        ok &= np.all(np.floor(qp / vox) == cell, 1) & np.all(np.floor(qm / vox) == cell, 1)
        sp, _, vp = ref.query64(qp)
        sm, _, vm = ref.query64(qm)
        ok &= vp & vm
        fd[:, k] = (sp - sm) / (2 * h)
    print(f"FD points kept: {ok.sum()} of {len(ok)}")
    assert ok.sum() > 5000
    Jg = J[valid][:30000]
    rel_fd = np.linalg.norm(fd[ok] - Jg[ok]) / np.linalg.norm(fd[ok])
    print(f"J vs central differences: {rel_fd:.2e}")
    assert rel_fd < 1e-5
    # pose gradient (FP64 reduction on the device)
    pg = d_pg.cpu().numpy()
    want = (up[valid, None].astype(np.float64) * J64[valid]).sum(0)
    assert np.linalg.norm(pg - want) / np.linalg.norm(want) < 1e-5
    # voxel gradient against the oracle's CPU scatter, voxel by voxel
    keys, _, _ = vol.export_blocks()
    nb = len(keys)
    vg = d_vg.cpu().numpy()[: nb * 512].reshape(nb, 512)
    keys, vg = oracle.sort_blocks(keys, vg)
    # the twin scatter forms the kernels' FP32 weights and products and sums them in FP64, so
    # only the device's FP32 accumulation differs: at most about k ulp of sum_i |up_i w_ik| for
    # k contributions; the FP64 scatter (pinned to finite differences) differs further by the
    # FP32 rounding of the weights, a few 1e-7 of sum_i |up_i| per touching point
    rk, g_twin, used = ref.scatter_grad(q, up, twin=True)
    _, g_ref, _ = ref.scatter_grad(q, up)
    _, g_abs, _ = ref.scatter_grad(q, np.abs(up), twin=True)
    assert np.array_equal(keys, rk) and used == valid.sum()
    err = np.abs(vg.astype(np.float64) - g_twin)
    err64 = np.abs(vg.astype(np.float64) - g_ref)
    print(f"voxel gradient: {int((g_abs > 0).sum())} voxels, max error {err.max():.2e} against the twin scatter, "
          f"{err64.max():.2e} against the FP64 scatter (max |g| {np.abs(g_ref).max():.2f})")
    assert np.all(err <= 1e-5 * g_abs + 1e-12)
    assert err64.max() < 1e-5 * np.abs(g_ref).max()


@pytest.mark.parametrize("scale", [1.0, 0.25])
def test_gauss_newton_on_the_fp16_volume(vh, wl, fp16_volume, scale):
    w, vol, ref = fp16_volume
    pts, T_true, xi = _held_out(vh, wl, w)
    T0 = T_true @ se3_exp(scale * xi)
    e0 = pose_error(T0, T_true)
    print(f"GN: {len(pts)} points, error x{scale}: {e0[0]:.3f} deg {1e3 * e0[1]:.2f} mm", flush=True)
    T1, st = vol.gn_refine(pts, T0, 10)
    Tc, trace = lm_reference(ref, pts, T0, 10, w["trunc"])
    errs = [pose_error(P, T_true) for P in st["pose"]]
    print("per evaluation: " + ", ".join(f"{a:.3f} deg {1e3 * b:.2f} mm" for a, b in errs))
    assert np.abs(T1 - Tc).max() < 1e-6
    for k in range(11):
        assert np.abs(np.vstack([st["pose"][k][:3], [0, 0, 0, 1]]) - trace[k]).max() < 1e-6, k
    e1 = errs[-1]
    assert e1[0] < 0.1 and e1[1] < 3e-3
    # no drift after the fourth evaluation: the error does not grow past its best so far
    best = min(e[1] for e in errs[:5])
    assert max(e[1] for e in errs[4:]) < best + 1e-3
