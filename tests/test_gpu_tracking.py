"""GPU parity of the tracked sequence (vh_bind_tracking): per frame, the device back-projects
the depth on a stride grid, refines the prior pose with the Levenberg-Marquardt solver against
the volume integrated so far, and integrates the frame at the refined pose.

Bars: every refined pose within 1e-6 (entry-wise) of the CPU FP64 restatement of the solver
(tests/gn_reference.py) run from the same prior on the oracle's volume of the preceding
frames; the first frame (empty volume) keeps its prior exactly; the per-evaluation stats
agree with the restatement's accepted poses; the final volume is bit-identical to the
oracle's after integrating the same frames at the device's refined poses; the refined pose
never costs more than the prior. (Whether it is closer to the true pose depends on the view:
on a map of a few frames of mostly planar room walls the objective can be lower at a pose
that slides along the planes than at the true one; the tracked benchmark reports accuracy.)
"""
import numpy as np
import pytest

from gn_reference import backproject_grid, lm_reference, pose_error, se3_exp, tracked_cost

pytestmark = pytest.mark.gpu


def noisy_priors(poses, sigma_t, sigma_deg, seed):
    """Frame 0 keeps its pose (it anchors the map); frame f >= 1 gets a camera-frame error
    exp(xi_f), xi_f ~ N(0, diag(sigma_t^2 I3, sigma_rot^2 I3))."""
    rng = np.random.default_rng(seed)
    out = np.asarray(poses, np.float64).copy()
    for f in range(1, len(out)):
        xi = np.concatenate([rng.normal(0, sigma_t, 3), rng.normal(0, np.deg2rad(sigma_deg), 3)])
        out[f] = out[f] @ se3_exp(xi)
    return out.astype(np.float32)


def odometry_increments(poses, sigma_t, sigma_deg, seed):
    """inc_0 = T_0; inc_f = (T_{f-1}^-1 T_f) exp(xi_f) for f >= 1 (a noisy camera-frame motion)."""
    rng = np.random.default_rng(seed)
    P = np.asarray(poses, np.float64)
    out = P.copy()
    for f in range(1, len(P)):
        xi = np.concatenate([rng.normal(0, sigma_t, 3), rng.normal(0, np.deg2rad(sigma_deg), 3)])
        out[f] = np.linalg.inv(P[f - 1]) @ P[f] @ se3_exp(xi)
    return out.astype(np.float32)


# (sigma_t, sigma_rot, depth_points) of the motion prior: the odometry's own noise, the depth
# image counted as 30 points
ODOMETRY_PRIOR = (0.005, float(np.deg2rad(0.5)), 30.0)


@pytest.mark.parametrize("storage,mode,carve,motion", [(0, 0, 0, None), (1, 0, 0, None), (0, 1, 0, None),
                                                      (0, 0, 1, None), (0, 1, 0, ODOMETRY_PRIOR),
                                                      (2, 1, 1, ODOMETRY_PRIOR)])
def test_tracked_sequence_matches_cpu_solver_and_oracle(vh, oracle, wl, small_frames, storage, mode, carve, motion):
    import torch

    w = dict(wl.WORKLOADS["room_small_1cm"])
    if storage:
        w.update(storage=storage, weight_cap=255)
    w["carve"] = carve
    depth, poses = small_frames
    nf, iters, stride = 6, 4, 4
    depth, poses = depth[:nf], poses[:nf]
    if mode == vh.PRIOR_INCREMENT:
        prior = odometry_increments(poses, 0.005, 0.5, wl.SEED + 2)
    else:
        prior = noisy_priors(poses, 0.01, 1.0, wl.SEED + 2)
    print(f"tracking: room_small_1cm storage {storage}, prior mode {mode}, carve {carve}, motion prior {motion}, "
          f"{nf} frames, {iters} iterations, stride {stride}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    d_depth = torch.from_numpy(np.ascontiguousarray(depth).view(np.int16)).cuda()
    d_prior = torch.from_numpy(prior.reshape(nf, 16)).cuda()
    d_out = torch.zeros(nf, 16, dtype=torch.float32, device="cuda")
    d_stats = torch.zeros(nf * (iters + 1) * vh.GN_STATS, dtype=torch.float64, device="cuda")
    st, sr, npts = motion or (0.0, 0.0, 0.0)
    vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, nf, d_prior, d_out, d_stats, iters, stride, prior_mode=mode,
                      prior_sigma_t=st, prior_sigma_rot=sr, depth_points=npts)
    vol.integrate_next(nf)
    vol.sync()
    out = d_out.cpu().numpy().reshape(nf, 4, 4)
    stats = d_stats.cpu().numpy().reshape(nf, iters + 1, vh.GN_STATS)

    assert np.array_equal(out[0], prior[0]), "first frame: the prior must be kept on an empty volume"
    ref = oracle.Oracle(w)
    errs = []
    for f in range(nf):
        pts = backproject_grid(depth[f], stride, 0.001, w["fx"], w["fy"], w["cx"], w["cy"], w["max_depth"])
        start = prior[f].astype(np.float64)
        if mode == vh.PRIOR_INCREMENT and f > 0:
            start = out[f - 1].astype(np.float64) @ start
        Tc, trace = lm_reference(ref, pts, start, iters, w["trunc"], prior=motion)
        dev = np.abs(out[f].astype(np.float64) - Tc).max()
        assert dev < 1e-6, (f, dev)
        for it in range(iters + 1):
            Ts = np.vstack([stats[f, it, 5:17].reshape(3, 4), [0, 0, 0, 1]])
            assert np.abs(Ts - trace[it]).max() < 1e-6, (f, it)
        if f == 0:
            assert stats[0, :, 0].max() == 0, "no point can have a sample in an empty volume"
        errs.append((pose_error(start, poses[f]), pose_error(out[f], poses[f])))
        # Levenberg-Marquardt never accepts a cost increase
        assert tracked_cost(ref, pts, Tc, w["trunc"], start, motion) <= tracked_cost(ref, pts, start, w["trunc"],
                                                                                    start, motion)
        ref.integrate(depth[f], out[f])
    for f, (e0, e1) in enumerate(errs):
        print(f"frame {f}: start {e0[0]:.3f} deg {e0[1] * 1000:.2f} mm -> refined {e1[0]:.3f} deg "
              f"{e1[1] * 1000:.2f} mm, rms {stats[f, 0, 1]:.4f} -> {stats[f, -1, 1]:.4f}")
    keys, tsdf, wt = vol.export_blocks()
    keys, tsdf, wt = oracle.sort_blocks(keys, tsdf, wt)
    rk, rs, rw, _, _ = ref.export()
    assert np.array_equal(keys, rk)
    assert np.array_equal(tsdf.view(np.uint32), rs.view(np.uint32))
    assert np.array_equal(wt, rw)
    vol.close()


def test_tracking_rejects_bad_arguments(vh, wl):
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    vol = vh.Volume(vh.config_from_workload(w))
    d = torch.zeros(2, 480, 640, dtype=torch.int16, device="cuda")
    p = torch.zeros(2, 16, dtype=torch.float32, device="cuda")
    for iters, stride, mode in ((64, 4, 0), (5, 0, 0), (5, 65, 0), (5, 4, 2)):
        with pytest.raises(vh.VhError) as e:
            vol.bind_tracking(d, vh.DEPTH_U16, 0.001, 2, p, p, None, iters, stride, prior_mode=mode)
        assert e.value.code == vh.VH_ERR_INVALID_ARG
    for bad in (dict(prior_sigma_t=-1.0), dict(prior_sigma_rot=float("nan")), dict(depth_points=float("inf"))):
        with pytest.raises(vh.VhError) as e:
            vol.bind_tracking(d, vh.DEPTH_U16, 0.001, 2, p, p, None, 5, 4, **bad)
        assert e.value.code == vh.VH_ERR_INVALID_ARG and "depth_points" in str(e.value)
    vol.close()
