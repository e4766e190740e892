"""GPU check of vh_gn_refine_device_noalloc, the pose refinement whose solver scratch is the
library's static device array: on two calls in a row (the second with fewer points) it returns
exactly what vh_gn_refine returns on the same points (same kernels, same reduction order), it
rejects more iterations than its scratch holds, and a call after the first leaves the device's
free memory unchanged."""
import numpy as np
import pytest

from gn_reference import backproject_grid, se3_exp

pytestmark = pytest.mark.gpu


def test_noalloc_refine_equals_allocating_refine(vh, wl, small_frames):
    import torch

    w = wl.WORKLOADS["room_small_1cm"]
    depth, poses = small_frames
    nf = 10
    print(f"no-alloc refine: room_small_1cm, {nf} frames integrated, held-out frame {nf}", flush=True)
    vol = vh.Volume(vh.config_from_workload(w))
    for k in range(nf):
        vol.integrate(depth[k], poses[k])
    rng = np.random.default_rng(wl.SEED + 2)
    xi = np.concatenate([rng.normal(0, 0.01, 3), rng.normal(0, np.deg2rad(1.0), 3)])
    T0 = poses[nf].astype(np.float64) @ se3_exp(xi)
    pts_all = backproject_grid(depth[nf], 2, 0.001, w["fx"], w["fy"], w["cx"], w["cy"], w["max_depth"])
    # two calls in a row, the second with a smaller input (and fewer iterations)
    for n, iters in ((len(pts_all), 8), (len(pts_all) // 3, 5)):
        pts = np.ascontiguousarray(pts_all[:n])
        d_pts = torch.from_numpy(pts).cuda()
        Ta, sa = vol.gn_refine(pts, T0, iters)
        Tb, sb = vol.gn_refine_device_noalloc(d_pts, n, T0, iters)
        print(f"  {n} points, {iters} iterations: rms {sa['rms'][0]:.5f} -> {sa['rms'][-1]:.5f}", flush=True)
        assert np.array_equal(Ta, Tb)
        for key in ("n_valid", "rms", "step_norm"):
            assert sa[key] == sb[key], key
        for Pa, Pb in zip(sa["pose"], sb["pose"]):
            assert np.array_equal(Pa, Pb)
        assert sa["rms"][-1] <= sa["rms"][0]
    # a later call allocates nothing on the device
    d_pts = torch.from_numpy(np.ascontiguousarray(pts_all)).cuda()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        vol.gn_refine_device_noalloc(d_pts, len(pts_all), T0, 8)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    # sizes beyond the static scratch are refused with a clear error
    with pytest.raises(vh.VhError) as e:
        vol.gn_refine_device_noalloc(d_pts, len(pts_all), T0, 64)
    assert e.value.code == vh.VH_ERR_INVALID_ARG and "63" in str(e.value)
    vol.close()
