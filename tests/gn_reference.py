"""CPU FP64 restatement of the device pose solver, used by the GPU tests as their reference.

The device solver (vh_kernels.hip: k_gn_accum, k_gn_solve; driven by vh_gn_refine and by the
tracked-sequence graph) evaluates r_i = sdf(R p_i + t) with the trilinear sample, forms
J_i = [g_i, q_i x g_i] for a left se(3) perturbation, sums J^T J, J^T r and the cost in FP64
and takes Levenberg-Marquardt steps. This module recomputes the same from the oracle's FP32
twin of the sample (oracle.Oracle.query_twin) with numpy FP64 sums, so a GPU result can be
compared to it; the se(3) exponential is the closed form (Rodrigues with the V matrix).
"""
import numpy as np


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def se3_exp(xi):
    """exp of xi = (rho, phi) as a 4x4 (FP64)."""
    rho, phi = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(phi)
    K = skew(phi)
    if th < 1e-9:
        R, V = np.eye(3) + K, np.eye(3) + 0.5 * K
    else:
        R = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
        V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * K + (th - np.sin(th)) / th ** 3 * K @ K
    T = np.eye(4)
    T[:3, :3], T[:3, 3] = R, V @ rho
    return T


def world_points(T, pts):
    """q = R p + t in FP64 with the kernels' operation order ((x*T0 + y*T1) + z*T2) + T3, from
    FP32 points and pose entries."""
    T = np.asarray(np.asarray(T, np.float32), np.float64)
    p = np.asarray(np.asarray(pts, np.float32), np.float64)
    q = np.empty_like(p)
    for r in range(3):
        q[:, r] = ((p[:, 0] * T[r, 0] + p[:, 1] * T[r, 1]) + p[:, 2] * T[r, 2]) + T[r, 3]
    return q


def world_points_d(T, pts):
    """As world_points, with FP64 pose entries (the solver's candidates are FP64)."""
    T = np.asarray(T, np.float64)
    p = np.asarray(np.asarray(pts, np.float32), np.float64)
    q = np.empty_like(p)
    for r in range(3):
        q[:, r] = ((p[:, 0] * T[r, 0] + p[:, 1] * T[r, 1]) + p[:, 2] * T[r, 2]) + T[r, 3]
    return q


def lm_cost(ref, pts, T):
    """The solver's cost at pose T: sum r^2 over points with a valid sample + 1 per other point."""
    q = world_points_d(T, pts)
    s, _, v = ref.query_twin(q)
    return float((s[v].astype(np.float64) ** 2).sum() + (len(pts) - v.sum()))


def lm_reference(ref, pts, T0, iters, lam0=1e-4):
    """Levenberg-Marquardt as the device runs it: cost = sum r^2 + 1 per point without a
    valid sample; accept if the cost does not grow (lambda / 3, floor 1e-9), else reject
    (lambda * 4, ceiling 1e9); step (A + lambda diag(A) + 1e-12 I) x = -b, T <- exp(x^) T;
    iters steps and a last evaluation. Returns the accepted pose and the accepted pose after
    each evaluation."""
    n = len(pts)

    def evaluate(T):
        q = world_points_d(T, pts)
        s, g, v = ref.query_twin(q)
        g64 = g.astype(np.float64)
        J = np.concatenate([g64, np.cross(q, g64)], 1)[v]
        r = s.astype(np.float64)[v]
        return (r ** 2).sum() + (n - v.sum()), J.T @ J, J.T @ r

    cand, acc, lam, out = np.asarray(T0, np.float64).copy(), None, lam0, []
    for it in range(iters + 1):
        cost, A, b = evaluate(cand)
        if acc is None or cost <= acc[1]:
            lam = lam if acc is None else max(lam / 3, 1e-9)
            acc = (cand, cost, A, b)
        else:
            lam = min(lam * 4, 1e9)
        out.append(acc[0])
        if it == iters:
            break
        Ad = acc[2] + np.diag(lam * np.diag(acc[2]) + 1e-12)
        x = np.linalg.solve(Ad, -acc[3])
        cand = se3_exp(x) @ acc[0]
    return acc[0], out


def pose_error(T, T_true):
    """(rotation error in degrees, translation error in metres)."""
    T, T_true = np.asarray(T, np.float64), np.asarray(T_true, np.float64)
    D = T @ np.linalg.inv(T_true)
    return (float(np.degrees(np.arccos(np.clip((np.trace(D[:3, :3]) - 1) / 2, -1, 1)))),
            float(np.linalg.norm(T[:3, 3] - T_true[:3, 3])))


def backproject_grid(depth_u16, stride, scale, fx, fy, cx, cy, max_depth):
    """Camera-frame points of every stride-th pixel of every stride-th row with a valid depth
    (0 < d <= max_depth), row-major, in FP32 with the device's operation order:
    ((u - cx) * d) / fx, ((v - cy) * d) / fy, d; d = depth * scale."""
    f32 = np.float32
    d = depth_u16[::stride, ::stride].astype(f32) * f32(scale)
    v, u = np.mgrid[0:depth_u16.shape[0]:stride, 0:depth_u16.shape[1]:stride]
    ok = (d > 0) & (d <= f32(max_depth))
    x = ((u.astype(f32) - f32(cx)) * d) / f32(fx)
    y = ((v.astype(f32) - f32(cy)) * d) / f32(fy)
    return np.stack([x[ok], y[ok], d[ok]], 1).astype(f32)
