"""GPU parity of tracking at the large-floor workload (floor_large_4mm: 4 mm voxels, 1.6 cm
truncation, camera 1.5-3 m from the walls): the first 150 frames, each starting from the
refined pose of the previous frame times a noisy odometry increment (0.5 mm and 0.05 degrees
per axis per frame, seeded), refined by 5 solver iterations on every 4th pixel up to 3 m with
the odometry's noise as the motion prior and the depth term counted as 3 points, then
integrated at the refined pose.

Bars: every refined pose within 1e-6 (entry-wise) of the CPU FP64 restatement of the solver
(tests/gn_reference.py) run from the same start on the oracle's volume of the preceding frames;
the final volume bit-identical to the oracle's after integrating the frames at the device's
poses; the tracked poses closer to the truth than dead reckoning (mean rotation and
translation error over the frames).
"""
import numpy as np
import pytest

from gn_reference import backproject_grid, lm_reference, pose_error
from test_gpu_tracking import odometry_increments

pytestmark = pytest.mark.gpu

NF, ITERS, STRIDE, ZMAX = 150, 5, 4, 3.0
SIGMA_T, SIGMA_DEG, DEPTH_POINTS = 0.0005, 0.05, 3.0


def test_floor_large_tracked_subset_matches_cpu_solver_and_oracle(vh, oracle, wl):
    import torch

    w = dict(wl.WORKLOADS["floor_large_4mm"])
    scene = vh.Scene(w["scene"], wl.SEED)
    depth, poses = scene.frames(0, NF)
    inc = odometry_increments(poses, SIGMA_T, SIGMA_DEG, wl.SEED + 2)
    prior = (SIGMA_T, float(np.deg2rad(SIGMA_DEG)), DEPTH_POINTS)
    print(f"tracking: floor_large_4mm {NF} frames, increments, motion prior {prior}, {ITERS} iterations, "
          f"stride {STRIDE}, points up to {ZMAX} m", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    d_depth = torch.from_numpy(np.ascontiguousarray(depth).view(np.int16)).cuda()
    d_inc = torch.from_numpy(inc.reshape(NF, 16)).cuda()
    d_out = torch.zeros(NF, 16, dtype=torch.float32, device="cuda")
    vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, NF, d_inc, d_out, None, ITERS, STRIDE, ZMAX,
                      vh.PRIOR_INCREMENT, prior_sigma_t=prior[0], prior_sigma_rot=prior[1], depth_points=prior[2])
    vol.integrate_next(NF)
    vol.sync()
    out = d_out.cpu().numpy().reshape(NF, 4, 4)
    print("device done; CPU restatement per frame", flush=True)

    assert np.array_equal(out[0], inc[0])
    ref = oracle.Oracle(w)
    worst = 0.0
    for f in range(NF):
        if f > 0:
            pts = backproject_grid(depth[f], STRIDE, 0.001, w["fx"], w["fy"], w["cx"], w["cy"], ZMAX)
            start = out[f - 1].astype(np.float64) @ inc[f].astype(np.float64)
            Tc, _ = lm_reference(ref, pts, start, ITERS, w["trunc"], prior=prior)
            dev = np.abs(out[f].astype(np.float64) - Tc).max()
            worst = max(worst, dev)
            assert dev < 1e-6, (f, dev)
        ref.integrate(depth[f], out[f])
    print(f"largest pose deviation {worst:.2e}", flush=True)

    keys, tsdf, wt = vol.export_blocks()
    keys, tsdf, wt = oracle.sort_blocks(keys, tsdf, wt)
    rk, rs, rw, _, _ = ref.export()
    assert np.array_equal(keys, rk)
    assert np.array_equal(wt, rw)
    assert np.array_equal(tsdf.view(np.uint32), rs.view(np.uint32))
    vol.close()

    dead = [inc[0].astype(np.float64)]
    for f in range(1, NF):
        dead.append(dead[-1] @ inc[f].astype(np.float64))
    e_dead = np.array([pose_error(dead[f], poses[f]) for f in range(1, NF)])
    e_trk = np.array([pose_error(out[f], poses[f]) for f in range(1, NF)])
    print(f"mean error: dead reckoning {e_dead[:, 0].mean():.3f} deg {1e3 * e_dead[:, 1].mean():.1f} mm, "
          f"tracked {e_trk[:, 0].mean():.3f} deg {1e3 * e_trk[:, 1].mean():.1f} mm", flush=True)
    assert e_trk[:, 0].mean() < e_dead[:, 0].mean() and e_trk[:, 1].mean() < e_dead[:, 1].mean()
