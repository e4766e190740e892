"""GPU parity of the trilinear SDF query, its backward pass and Gauss-Newton refinement.

Bars: query values bit-identical to the oracle's FP32 twin; per-point pose Jacobians within
1e-5 (norm-wise relative) of the oracle's FP64 analytic gradient at the same points and of
central finite differences (h = 1e-4, FP64) on points whose stencil stays inside one lattice
cell; aggregate pose gradients within 1e-5 of FD relative to the magnitude of their terms;
Gauss-Newton recovering a seeded 2 deg / 3 cm pose error.
"""
import numpy as np
import pytest

from conftest import look_down_z, plane_depth
from gn_reference import lm_reference, pose_error, se3_exp, skew, world_points  # noqa: F401

pytestmark = pytest.mark.gpu


def backproject(depth_m, T, step=2, fx=525.0, fy=525.0, cx=319.5, cy=239.5):
    """Camera-frame points of every step-th valid pixel."""
    v, u = np.nonzero(depth_m[::step, ::step] > 0)
    v, u = v * step, u * step
    z = depth_m[v, u].astype(np.float64)
    return np.stack([(u - cx) * z / fx, (v - cy) * z / fy, z], 1).astype(np.float32)


@pytest.fixture(scope="module")
def room(vh, oracle, wl, small_frames):
    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    vol = vh.Volume(vh.config_from_workload(w))
    ref = oracle.Oracle(w)
    for k in range(len(depth)):
        vol.integrate(depth[k], poses[k])
        ref.integrate(depth[k], poses[k])
    yield w, vol, ref
    vol.close()


def surface_points(room, n, seed):
    w, vol, ref = room
    xyz, _ = vol.extract_mesh()
    rng = np.random.default_rng(seed)
    sel = rng.choice(len(xyz), n)
    p = xyz[sel].reshape(-1, 3, 3).mean(1)
    return (p + rng.normal(0, 0.01, p.shape)).astype(np.float32)


def test_query_bit_exact_vs_oracle(room):
    w, vol, ref = room
    pts = surface_points(room, 200_000, 0)
    print("query: 200k points near the room_small_1cm surface", flush=True)
    sdf, valid = vol.query(pts)
    s32, s64, g64, rvalid = ref.query(pts)
    assert np.array_equal(valid, rvalid)
    assert valid.mean() > 0.5
    assert np.array_equal(sdf[valid].view(np.uint32), s32[valid].view(np.uint32))
    assert np.abs(sdf[valid] - s64[valid]).max() < 5e-6
    assert np.all(sdf[~valid] == 0)


def test_query_affine_field_and_lattice_points(vh, oracle, wl):
    """A fronto-parallel plane gives an affine TSDF (D - z)/mu inside the band; trilinear
    interpolation reproduces it exactly between lattice points."""
    w = wl.WORKLOADS["room_small_1cm"]
    vol = vh.Volume(vh.config_from_workload(w))
    T = look_down_z(0.0, 0.0, 0.0)
    vol.integrate(plane_depth(w, 1500), T)
    rng = np.random.default_rng(3)
    z = rng.uniform(1.475, 1.525, 5000)
    x, y = rng.uniform(-0.3, 0.3, 5000), rng.uniform(-0.2, 0.2, 5000)
    pts = np.stack([x, y, z], 1).astype(np.float32)
    sdf, valid = vol.query(pts)
    assert valid.all()
    d = 1500 * np.float64(np.float32(0.001))
    want = (d - pts[:, 2].astype(np.float64)) / 0.04
    assert np.abs(sdf - want).max() < 2e-5
    # at lattice points the query returns the stored voxel value
    keys, tsdf, wt = vol.export_blocks()
    b = int(np.argmax((wt > 0).sum(1)))
    i = int(np.nonzero(wt[b] > 0)[0][0])
    g = keys[b] * 8 + np.array([i >> 6, (i >> 3) & 7, i & 7])
    s1, v1 = vol.query((g * np.float32(0.01)).astype(np.float32)[None])
    if v1[0]:
        assert abs(s1[0] - tsdf[b, i]) < 1e-5
    vol.close()


def test_backward_jacobian_and_fd(room, oracle):
    import torch

    w, vol, ref = room
    rng = np.random.default_rng(5)
    T = np.eye(4, dtype=np.float32)  # identity pose: world points = camera points
    pts = surface_points(room, 100_000, 6)
    n = len(pts)
    d_pts = torch.from_numpy(pts).cuda()
    up = rng.normal(size=n).astype(np.float32)
    d_up = torch.from_numpy(up).cuda()
    d_r = torch.zeros(n, dtype=torch.float32, device="cuda")
    d_valid = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_J = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    d_pg = torch.zeros(6, dtype=torch.float64, device="cuda")
    pool_vox, _, nb = vol.pool_view()
    d_vg = torch.zeros(w["pool_blocks"] * 512, dtype=torch.float32, device="cuda")
    print("backward: 100k points, identity pose", flush=True)
    vol.backward_device(d_pts, n, T, d_up, d_r, d_valid, d_J, d_pg, d_vg)
    torch.cuda.synchronize()
    r, valid, J, pg = d_r.cpu().numpy(), d_valid.cpu().numpy().astype(bool), d_J.cpu().numpy(), d_pg.cpu().numpy()
    q = world_points(T, pts)
    s64, g64, v64 = ref.query64(q)
    assert np.array_equal(valid, v64)
    J64 = np.concatenate([g64, np.cross(q, g64)], 1)
    rel = np.linalg.norm(J[valid] - J64[valid]) / np.linalg.norm(J64[valid])
    print(f"J relative error vs FP64 analytic: {rel:.2e}")
    assert rel < 1e-5
    assert np.abs(r[valid] - s64[valid]).max() < 5e-6
    # aggregate pose gradient (FP64 reduction on the GPU)
    want = (up[valid, None].astype(np.float64) * J64[valid]).sum(0)
    assert np.linalg.norm(pg - want) / np.linalg.norm(want) < 1e-5
    # voxel gradient: trilinear weights sum to one
    vg = d_vg.cpu().numpy()
    assert abs(vg.sum() - up[valid].astype(np.float64).sum()) < 1e-3 * max(1.0, np.abs(up[valid]).sum() / 1e3)
    # FP64 analytic gradient vs central differences, points whose stencil stays in one cell
    h, vox = 1e-4, float(np.float32(0.01))
    q64 = q[valid][:20000]
    cell = np.floor(q64 / vox)
    ok = np.ones(len(q64), bool)
    fd = np.zeros((len(q64), 6))
    for k in range(6):
        xi = np.zeros(6)
        xi[k] = h
        qp = (se3_exp(xi)[:3, :3] @ q64.T).T + se3_exp(xi)[:3, 3]
        qm = (se3_exp(-xi)[:3, :3] @ q64.T).T + se3_exp(-xi)[:3, 3]
        ok &= np.all(np.floor(qp / vox) == cell, 1) & np.all(np.floor(qm / vox) == cell, 1)
        sp, _, vp = ref.query64(qp)
        sm, _, vm = ref.query64(qm)
        ok &= vp & vm
        fd[:, k] = (sp - sm) / (2 * h)
    J64v = J64[valid][:20000]
    print(f"FD points kept: {ok.sum()} of {len(ok)}")
    assert ok.sum() > 5000
    rel_fd = np.linalg.norm(fd[ok] - J64v[ok]) / np.linalg.norm(J64v[ok])
    print(f"analytic vs FD relative error: {rel_fd:.2e}")
    assert rel_fd < 1e-5
    Jg = J[valid][:20000]
    rel_gpu_fd = np.linalg.norm(fd[ok] - Jg[ok]) / np.linalg.norm(fd[ok])
    assert rel_gpu_fd < 1e-5
    agg_fd = (up[valid][:20000][ok, None] * fd[ok]).sum(0)
    agg_gpu = (up[valid][:20000][ok, None].astype(np.float64) * Jg[ok]).sum(0)
    assert np.linalg.norm(agg_gpu - agg_fd) / np.linalg.norm(agg_fd) < 1e-5


class TrilinearSDF:
    """torch.autograd.Function over vh_query_sdf_backward_device (test harness only):
    r = sdf(exp(xi^) T p); gradients w.r.t. xi (6, FP64) and the pool's voxel tsdf values."""

    @staticmethod
    def make(vol, pool_voxels):
        import torch

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, xi, vox_dummy, pts, T):
                Tn = (se3_exp(xi.detach().cpu().numpy()) @ T).astype(np.float32)
                n = pts.shape[0]
                r = torch.zeros(n, dtype=torch.float32, device="cuda")
                valid = torch.zeros(n, dtype=torch.uint8, device="cuda")
                vol.backward_device(pts, n, Tn, None, r, valid)
                ctx.save_for_backward(pts)
                ctx.Tn = Tn
                return r

            @staticmethod
            def backward(ctx, grad_r):
                (pts,) = ctx.saved_tensors
                n = pts.shape[0]
                pg = torch.zeros(6, dtype=torch.float64, device="cuda")
                vg = torch.zeros(pool_voxels, dtype=torch.float32, device="cuda")
                vol.backward_device(pts, n, ctx.Tn, grad_r.contiguous().float(), None, None, None, pg, vg)
                return pg, vg, None, None

        return Fn


def test_autograd_harness_matches_fd(room, oracle):
    """loss = 1/2 sum_i (m_i r_i)^2 through the autograd harness, m_i selecting the points
    whose +-h stencils stay inside one lattice cell (where r is smooth in xi): the pose
    gradient matches central differences of the FP64 objective within 1e-5, and the voxel
    gradient sums to sum_i m_i r_i (trilinear weights sum to one)."""
    import torch

    w, vol, ref = room
    pts = surface_points(room, 50_000, 8)
    T = np.eye(4)
    q = world_points(T, pts)
    h, vox = 1e-4, float(np.float32(0.01))
    cell = np.floor(q / vox)
    _, _, m = ref.query64(q)
    stencil = []
    for k in range(6):
        for sgn in (1, -1):
            e = np.zeros(6)
            e[k] = sgn * h
            E = se3_exp(e)
            qs = (E[:3, :3] @ q.T).T + E[:3, 3]
            _, _, vs = ref.query64(qs)
            m &= vs & np.all(np.floor(qs / vox) == cell, 1)
            stencil.append(qs)
    print(f"autograd harness: {m.sum()} of {len(m)} points with in-cell stencils", flush=True)
    assert m.sum() > 10000
    d_pts = torch.from_numpy(pts).cuda()
    d_mask = torch.from_numpy(m.astype(np.float32)).cuda()
    Fn = TrilinearSDF.make(vol, w["pool_blocks"] * 512)
    xi = torch.zeros(6, dtype=torch.float64, device="cuda", requires_grad=True)
    vox_t = torch.zeros(w["pool_blocks"] * 512, dtype=torch.float32, device="cuda", requires_grad=True)
    r = Fn.apply(xi, vox_t, d_pts, T)
    loss = 0.5 * ((r * d_mask).double() ** 2).sum()
    loss.backward()
    g = xi.grad.cpu().numpy()
    fd = np.zeros(6)
    for k in range(6):
        sp, _, _ = ref.query64(stencil[2 * k])
        sm, _, _ = ref.query64(stencil[2 * k + 1])
        fd[k] = (0.5 * (sp[m] ** 2).sum() - 0.5 * (sm[m] ** 2).sum()) / (2 * h)
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    # sum_i r_i J_i cancels (r changes sign across the surface): compare the error with the
    # magnitude of the terms as well as with the sum
    s64, g64, _ = ref.query64(q)
    terms = np.abs(s64[m, None] * np.concatenate([g64[m], np.cross(q[m], g64[m])], 1)).sum(0)
    rel_terms = np.linalg.norm(g - fd) / np.linalg.norm(terms)
    print(f"pose gradient vs FD: relative error {rel:.2e} (sum), {rel_terms:.2e} (terms); "
          f"cancellation |sum|/sum|terms| = {np.linalg.norm(fd) / np.linalg.norm(terms):.3f}")
    assert rel_terms < 1e-5
    rm = (r * d_mask).double().sum().item()
    assert abs(vox_t.grad.double().sum().item() - rm) < 1e-4 * max(1.0, abs(rm))


def corner_view(pos, yaw_deg, pitch_deg):
    """Camera at pos looking along (yaw, pitch), image y pointing down (as the generator)."""
    yaw, pitch = np.radians(yaw_deg), np.radians(pitch_deg)
    f = np.array([np.cos(yaw) * np.cos(pitch), np.sin(yaw) * np.cos(pitch), np.sin(pitch)])
    r = np.array([np.sin(yaw), -np.cos(yaw), 0.0])
    M = np.eye(4)
    M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = r, np.cross(f, r), f, pos
    return M.astype(np.float32).astype(np.float64)


def test_gauss_newton_matches_cpu_reference_and_converges(vh, oracle, wl, room):
    """Held-out corner view of the small room (floor + two walls) with a seeded camera-frame
    pose error; the GPU's FP64 Levenberg-Marquardt (J^T J, J^T r, cost and the 6x6 solve on
    the device) follows the CPU FP64 restatement and reduces the error."""
    w, vol, ref = room
    scene = vh.Scene(w["scene"], wl.SEED)
    T_true = corner_view([2.0, 1.5, 1.5], 30.0, -25.0)
    dm = scene.clean_depth(1000, T_true.astype(np.float32))
    pts = backproject(dm, T_true, step=2)
    rng = np.random.default_rng(wl.SEED + 2)
    xi_err = np.concatenate([rng.normal(0, 0.03, 3), rng.normal(0, np.deg2rad(2.0), 3)]) * 0.25
    T0 = T_true @ se3_exp(xi_err)  # error in the camera frame
    e0 = pose_error(T0, T_true)
    print(f"GN: {len(pts)} points, initial error {e0[0]:.3f} deg {e0[1] * 1000:.2f} mm", flush=True)
    T1, st = vol.gn_refine(pts, T0, 10)
    Tc, _ = lm_reference(ref, pts, T0, 10, w["trunc"])
    e1, ec = pose_error(T1, T_true), pose_error(Tc, T_true)
    print(f"GPU: {e1[0]:.4f} deg {e1[1] * 1000:.3f} mm;  CPU reference: {ec[0]:.4f} deg {ec[1] * 1000:.3f} mm")
    print(f"rms per evaluation {np.round(st['rms'], 5)}")
    assert np.abs(T1 - Tc).max() < 1e-6
    # residual RMS at the true pose: the floor of what the objective can reach
    s_true, _, v_true = ref.query64(world_points(T_true, pts))
    rms_true = float(np.sqrt((s_true[v_true] ** 2).mean()))
    print(f"rms at the true pose {rms_true:.5f}")
    assert st["rms"][-1] < 1.5 * rms_true
    assert e1[0] < 0.05 and e1[1] < e0[1]


