"""CPU FP64 restatement of the device pose solver, used by the GPU tests as their reference.

The device solver (vh_kernels.hip: k_gn_accum, k_gn_solve; driven by vh_gn_refine and by the
tracked-sequence graph) evaluates r_i = sdf(R p_i + t) with the trilinear sample, forms
J_i = [g_i, q_i x g_i] for a left se(3) perturbation, and takes Levenberg-Marquardt steps on a
robust, noise-weighted cost: a point measured at depth z has the axial noise
sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m of structured-light sensors (Nguyen, Izadi and
Lovell 2012), its residual is scaled to e = r mu / sigma(z), and the cost is the Huber loss
(threshold HUBER) of e; each step solves the iteratively reweighted normal equations
(J^T W J, J^T W r with W = psi(e)/e * s^2). This module recomputes the same from the oracle's
FP32 twin of the sample (oracle.Oracle.query_twin) with numpy FP64 sums, so a GPU result can be
compared to it; the se(3) exponential is the closed form (Rodrigues with the V matrix).
"""
import numpy as np

HUBER = 2.0  # threshold of the Huber loss, in units of the depth noise


def noise_scale(pts, mu):
    """s = mu / sigma(z) per point (z = the point's camera-frame depth, FP32 widened)."""
    z = np.asarray(np.asarray(pts, np.float32)[:, 2], np.float64)
    return np.float64(np.float32(mu)) / (0.0012 + 0.0019 * ((z - 0.4) * (z - 0.4)))


def huber_rho(e):
    a = np.abs(e)
    return np.where(a <= HUBER, 0.5 * e * e, HUBER * (a - 0.5 * HUBER))


def huber_weight(e):
    a = np.abs(e)
    return np.where(a <= HUBER, 1.0, HUBER / np.maximum(a, 1e-300))


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def se3_exp(xi):
    """exp of xi = (rho, phi) as a 4x4 (FP64)."""
    rho, phi = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th2 = float(phi @ phi)
    th = np.sqrt(th2)
    K = skew(phi)
    if th < 1e-5:  # series of sin/th, (1 - cos)/th^2, (th - sin)/th^3 (the closed forms cancel)
        a, b, c = 1.0 - th2 / 6.0, 0.5 - th2 / 24.0, 1.0 / 6.0 - th2 / 120.0
    else:
        a, b, c = np.sin(th) / th, (1 - np.cos(th)) / th2, (th - np.sin(th)) / (th2 * th)
    R = np.eye(3) + a * K + b * K @ K
    V = np.eye(3) + b * K + c * K @ K
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


def _evaluate(ref, pts, T, sc):
    """(cost, J^T W J, J^T W r) at pose T."""
    q = world_points_d(T, pts)
    s, g, v = ref.query_twin(q)
    g64 = g.astype(np.float64)
    J = np.concatenate([g64, np.cross(q, g64)], 1)[v]
    r = s.astype(np.float64)[v]
    e = sc[v] * r
    W = huber_weight(e) * sc[v] * sc[v]
    cost = huber_rho(e).sum() + huber_rho(sc[~v]).sum()
    return cost, (J * W[:, None]).T @ J, (J * W[:, None]).T @ r


def lm_cost(ref, pts, T, mu):
    """The solver's cost at pose T: Huber loss of the noise-scaled residuals of the points with
    a valid sample, and of a saturated residual (r = 1) for the others."""
    return float(_evaluate(ref, pts, T, noise_scale(pts, mu))[0])


def se3_log(T):
    """log of a 4x4 rigid transform as xi = (rho, phi) (FP64), the inverse of se3_exp: the angle
    from atan2(|w|/2, (tr R - 1)/2) with w the skew part of R, phi = theta / (2 sin theta) w,
    rho = V^-1 t with V^-1 = I - K/2 + (1/theta^2 - (1 + cos theta)/(2 theta sin theta)) K^2."""
    T = np.asarray(T, np.float64)
    R, t = T[:3, :3], T[:3, 3]
    w = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    sn, cs = 0.5 * np.linalg.norm(w), 0.5 * (np.trace(R) - 1.0)
    th = np.arctan2(sn, cs)
    if th < 1e-5:
        f, c2 = 0.5 + th * th / 12.0, 1.0 / 12.0 + th * th / 720.0
    else:
        f, c2 = th / (2.0 * sn), 1.0 / (th * th) - (1.0 + cs) / (2.0 * th * sn)
    phi = f * w
    K = skew(phi)
    return np.concatenate([(np.eye(3) - 0.5 * K + c2 * K @ K) @ t, phi])


def adjoint(T):
    """Ad(T) on xi = (rho, phi): exp((Ad(T) xi)^) = T exp(xi^) T^-1."""
    T = np.asarray(T, np.float64)
    R, t = T[:3, :3], T[:3, 3]
    A = np.zeros((6, 6))
    A[:3, :3], A[:3, 3:], A[3:, 3:] = R, skew(t) @ R, R
    return A


def motion_prior(T_start, T, sigma_t, sigma_rot):
    """(cost, A, b) of the tracked frame's motion prior at pose T: xi = log(T_start^-1 T),
    cost = xi^T Lam xi / 2 with Lam = diag(I3 / sigma_t^2, I3 / sigma_rot^2) (a sigma of 0
    drops its half), A = M^T Lam M and b = M^T Lam xi with M = Ad(T^-1) (the Jacobian of xi
    under T <- exp(delta^) T at xi = 0). sigma_* as the device holds them: FP32 values."""
    st, sr = float(np.float32(sigma_t)), float(np.float32(sigma_rot))
    lt, lr = (1.0 / (st * st) if st > 0 else 0.0), (1.0 / (sr * sr) if sr > 0 else 0.0)
    lam = np.array([lt] * 3 + [lr] * 3)
    T = np.asarray(T, np.float64)
    xi = se3_log(np.linalg.inv(np.asarray(T_start, np.float64)) @ T)
    # M^T Lam M and M^T Lam xi in closed form (R R^T = I): [[lt I, -lt [t]x], [lt [t]x,
    # lt (|t|^2 I - t t^T) + lr I]] and (g, t x g + lr R xi_phi), g = lt R xi_rho
    R, t = T[:3, :3], T[:3, 3]
    K = skew(t)
    A = np.zeros((6, 6))
    A[:3, :3], A[:3, 3:], A[3:, :3] = lt * np.eye(3), -lt * K, lt * K
    A[3:, 3:] = lt * (float(t @ t) * np.eye(3) - np.outer(t, t)) + lr * np.eye(3)
    g = lt * (R @ xi[:3])
    b = np.concatenate([g, np.cross(t, g) + lr * (R @ xi[3:])])
    return 0.5 * float(xi @ (lam * xi)), A, b


def _tracked(ref, pts, T, sc, T_start, prior):
    cost, A, b = _evaluate(ref, pts, T, sc)
    if prior is None:
        return cost, A, b
    st, sr, n_eff = prior
    n_eff = float(np.float32(n_eff))
    if n_eff > 0 and len(pts):
        al = n_eff / float(len(pts))
        cost, A, b = al * cost, al * A, al * b
    if st > 0 or sr > 0:
        pc, pA, pb = motion_prior(T_start, T, st, sr)
        cost, A, b = cost + pc, A + pA, b + pb
    return cost, A, b


def tracked_cost(ref, pts, T, mu, T_start, prior):
    """The tracked frame's solver cost at pose T (lm_reference with this prior and start)."""
    return float(_tracked(ref, pts, T, noise_scale(pts, mu), np.asarray(T_start, np.float64), prior)[0])


def lm_reference(ref, pts, T0, iters, mu, lam0=1e-4, prior=None):
    """Levenberg-Marquardt as the device runs it on the robust cost of lm_cost: accept if the
    cost does not grow (lambda / 3, floor 1e-9), else reject (lambda * 4, ceiling 1e9); step
    (A + lambda diag(A) + 1e-12 I) x = -b with A = J^T W J and b = J^T W r at the accepted
    pose, T <- exp(x^) T; iters steps and a last evaluation. mu: the volume's truncation
    (metres). prior: None, or (sigma_t, sigma_rot, depth_points) of a tracked frame: the depth
    term (cost, A, b) scaled by depth_points / len(pts) when depth_points > 0, plus
    motion_prior centred on T0. Returns the accepted pose and the accepted pose after each
    evaluation."""
    sc = noise_scale(pts, mu)
    T_start = np.asarray(T0, np.float64).copy()

    def evaluate(T):
        return _tracked(ref, pts, T, sc, T_start, prior)

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
