#!/usr/bin/env python3
"""Benchmark: integrated frames/s and marching-cubes blocks/s on one MI355X.

Workload (default room_medium_5mm): 500 synthetic 640x480 uint16 depth frames of an
8x6x3 m room with 12 primitives, 5 mm voxels, 2 cm truncation, 2^21 buckets, 1.5M-block
pool, FP32 voxels; per frame one captured HIP graph launch integrates frame f and allocates
frame f+1 (the first allocation and the last integration of a call run alone); marching
cubes runs every 100 frames (5 extractions per pass, the last one final). The integration
kernel's time for the roofline comes from a separate pass in profiling mode, where
allocation and integration run as their own launches with HIP events around each.
One step = one pass over the whole sequence from an empty volume; the depth frames and
poses are resident in HBM before timing starts.

Reported beside the headline FP32 figure: the same pass with binary16 TSDF + uint16 weight
storage, the FP16/FP32 speedup and the FP16 error against the oracle's FP64 shadow, the
integration kernel's achieved HBM bandwidth (algorithmic bytes / kernel time, measured with
HIP events on the volume's stream), and the oracle's CPU throughput on a bounded sample.
Two further blocks: "gn" (pose refinement on the FP16 volume: 10 Levenberg-Marquardt
iterations over the ~300k back-projected pixels of a held-out frame from a seeded
2 deg / 3 cm pose error, the pose error after every evaluation) and "tracked" (the
floor_large_4mm workload: 2000 frames of a 20x20x4 m floor plan, 4 mm voxels, FP16, each
frame's pose - the previous refined pose times a noisy odometry increment - refined by 5
solver iterations against the volume before it is integrated, final marching cubes).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "frames/s integrated + MC blocks/s, FP32 vs FP16-store at ≤1e-3 TSDF err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def se3_exp(xi):
    """exp of xi = (rho, phi) in se(3) as a 4x4 (FP64 closed form)."""
    rho, phi = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = np.linalg.norm(phi)
    K = np.array([[0, -phi[2], phi[1]], [phi[2], 0, -phi[0]], [-phi[1], phi[0], 0]])
    if th < 1e-9:
        R, V = np.eye(3) + K, np.eye(3) + 0.5 * K
    else:
        R = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
        V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * K + (th - np.sin(th)) / th ** 3 * K @ K
    T = np.eye(4)
    T[:3, :3], T[:3, 3] = R, V @ rho
    return T


def pose_err(T, T_true):
    """(rotation error in degrees, translation error in millimetres)."""
    T, T_true = np.asarray(T, np.float64), np.asarray(T_true, np.float64)
    D = T[:3, :3] @ T_true[:3, :3].T
    return (float(np.degrees(np.arccos(np.clip((np.trace(D) - 1) / 2, -1, 1)))),
            float(1e3 * np.linalg.norm(T[:3, 3] - T_true[:3, 3])))


def backproject(depth_u16, w, stride=1, scale=0.001, zmax=None):
    """Camera-frame FP32 points ((u - cx) d / fx, (v - cy) d / fy, d) of the pixels with
    0 < d <= zmax (default max_depth)."""
    f32 = np.float32
    d = depth_u16[::stride, ::stride].astype(f32) * f32(scale)
    v, u = np.mgrid[0:depth_u16.shape[0]:stride, 0:depth_u16.shape[1]:stride]
    ok = (d > 0) & (d <= f32(zmax or w["max_depth"]))
    x = ((u.astype(f32) - f32(w["cx"])) * d) / f32(w["fx"])
    y = ((v.astype(f32) - f32(w["cy"])) * d) / f32(w["fy"])
    return np.stack([x[ok], y[ok], d[ok]], 1).astype(f32)


def gn_block(vh, wl, w, vol, nf):
    """Pose refinement on the FP16 volume after the full sequence: the held-out frame nf + 17
    of the camera path (never integrated), all its valid pixels as points (and, separately,
    those up to 3 m: far surfaces of a TSDF fused from noisy depth sit a few mm off, which
    biases the objective), a seeded camera-frame pose error xi ~ N(0, diag(3 cm, 2 deg)) (and
    the same error scaled by 1/4), 10 Levenberg-Marquardt iterations; pose error after every
    evaluation."""
    import torch

    scene = vh.Scene(w["scene"], wl.SEED)
    f = nf + 17
    d, T = scene.frames(f, 1)
    T_true = T[0].astype(np.float64)
    rng = np.random.default_rng(wl.SEED + 2)
    xi = np.concatenate([rng.normal(0, 0.03, 3), rng.normal(0, np.deg2rad(2.0), 3)])
    out = {"frame": f, "iterations": 10, "storage": "fp16 (binary16 + uint16 weight)",
           "perturbation": "camera frame, rho ~ N(0, 3 cm) and phi ~ N(0, 2 deg) per axis, seed SEED+2", "runs": []}
    for scale, zmax in ((1.0, None), (0.25, None), (1.0, 3.0), (0.25, 3.0)):
        pts = backproject(d[0], w, zmax=zmax)
        T0 = T_true @ se3_exp(scale * xi)
        vol.gn_refine(pts, T0, 10)  # warm (scratch allocation, code paths)
        torch.cuda.synchronize()
        t = time.perf_counter()
        T1, st = vol.gn_refine(pts, T0, 10)
        dt = time.perf_counter() - t
        errs = [pose_err(P, T_true) for P in st["pose"]]
        out["runs"].append({"error_scale": scale, "max_point_depth_m": zmax or w["max_depth"], "points": int(len(pts)),
                            "initial_err_deg_mm": [round(x, 4) for x in pose_err(T0, T_true)],
                            "err_deg_mm_per_eval": [[round(a, 4), round(b, 3)] for a, b in errs],
                            "rms_per_eval": [round(r, 5) for r in st["rms"]],
                            "n_valid_per_eval": st["n_valid"], "wall_ms": round(1e3 * dt, 3)})
        log(f"bench: gn x{scale} z<={zmax or w['max_depth']}: {len(pts)} points, {pose_err(T0, T_true)} -> {errs[-1]} "
            f"in {1e3 * dt:.2f} ms")
    return out


def tracked_block(vh, wl, n_frames, gn_iters=5, stride=4, zmax=3.0, sigma_t=0.0005, sigma_deg=0.05,
                  depth_points=3.0):
    """floor_large_4mm with tracking from odometry: frame 0 at its true pose; frame f > 0
    starts from the refined pose of frame f-1 times a noisy motion increment
    inc_f = (T_{f-1}^-1 T_f) exp(xi_f), xi_f ~ N(0, sigma_t) and N(0, sigma_deg) per axis
    (camera frame, seeded), and is refined by gn_iters solver iterations on every stride-th
    pixel up to zmax metres before integration, with the odometry's noise as the motion prior
    and the depth term counted as depth_points points. Dead reckoning (chaining the increments
    without refinement) and tracking without the motion prior are reported beside it."""
    import torch

    w = dict(wl.WORKLOADS["floor_large_4mm"])
    nf = n_frames or w["frames"]
    gn_iters = w.get("gn_iters", gn_iters)
    scene = vh.Scene(w["scene"], wl.SEED)
    t = time.perf_counter()
    depth, poses = scene.frames(0, nf)
    log(f"bench: tracked: generated {nf} floor_large_4mm frames in {time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(wl.SEED + 2)
    P = poses.astype(np.float64)
    inc = np.empty_like(P)
    inc[0] = P[0]
    for f in range(1, nf):
        xi = np.concatenate([rng.normal(0, sigma_t, 3), rng.normal(0, np.deg2rad(sigma_deg), 3)])
        inc[f] = np.linalg.inv(P[f - 1]) @ P[f] @ se3_exp(xi)
    inc = inc.astype(np.float32)
    dead = [inc[0].astype(np.float64)]
    for f in range(1, nf):
        dead.append(dead[-1] @ inc[f].astype(np.float64))
    d_depth = torch.from_numpy(depth.view(np.int16)).cuda()
    d_inc = torch.from_numpy(inc.reshape(nf, 16)).cuda()
    d_out = torch.zeros(nf, 16, dtype=torch.float32, device="cuda")
    d_stats = torch.zeros(nf * (gn_iters + 1) * vh.GN_STATS, dtype=torch.float64, device="cuda")
    del depth
    vol = vh.Volume(vh.config_from_workload(w))
    mode = vh.PRIOR_INCREMENT
    motion = dict(prior_sigma_t=sigma_t, prior_sigma_rot=float(np.deg2rad(sigma_deg)), depth_points=depth_points)
    res = {"workload": "floor_large_4mm", "frames": nf, "gn_iters_per_frame": gn_iters, "point_stride": stride,
           "max_point_depth_m": zmax, "prior": "refined pose of the previous frame x noisy odometry increment "
           f"(camera frame, N(0, {sigma_t * 1e3:g} mm) and N(0, {sigma_deg:g} deg) per axis per frame, seed SEED+2)",
           "motion_prior": {"sigma_t_m": sigma_t, "sigma_rot_deg": sigma_deg, "depth_points": depth_points}}
    summ = lambda e: {"mean": round(float(e.mean()), 4), "median": round(float(np.median(e)), 4),
                      "p95": round(float(np.percentile(e, 95)), 4), "max": round(float(e.max()), 4),
                      "last": round(float(e[-1]), 4)}
    try:
        # without the motion prior (errors only)
        vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, nf, d_inc, d_out, d_stats, gn_iters, stride, zmax, mode)
        vol.integrate_next(nf)
        vol.sync()
        out = d_out.cpu().numpy().reshape(nf, 4, 4)
        e_np = np.array([pose_err(out[f], P[f]) for f in range(1, nf)])
        res["without_motion_prior"] = {"rot_err_deg": summ(e_np[:, 0]), "trans_err_mm": summ(e_np[:, 1])}
        vol.reset()
        vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, nf, d_inc, d_out, d_stats, gn_iters, stride, zmax, mode,
                          **motion)
        vol.integrate_next(1)  # captures the per-frame graph
        vol.sync()
        vol.reset()
        vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, nf, d_inc, d_out, d_stats, gn_iters, stride, zmax, mode,
                          **motion)
        torch.cuda.synchronize()
        t = time.perf_counter()
        vol.integrate_next(nf)
        vol.sync()
        dt = time.perf_counter() - t
        st = vol.stats()
        cap = 64 << 20
        d_xyz = torch.empty(cap * 9, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        t = time.perf_counter()
        ntri = vol.extract_mesh_device(d_xyz, None, cap)
        mc_s = time.perf_counter() - t
        del d_xyz
        out = d_out.cpu().numpy().reshape(nf, 4, 4)
        e_dead = np.array([pose_err(dead[f], P[f]) for f in range(1, nf)])
        e_trk = np.array([pose_err(out[f], P[f]) for f in range(1, nf)])
        res.update({"fps": round(nf / dt, 2), "ms_per_frame": round(1e3 * dt / nf, 4),
                    "blocks": st["blocks_allocated"], "pool_blocks": w["pool_blocks"],
                    "mc_triangles": ntri, "mc_ms": round(1e3 * mc_s, 3),
                    "mc_blocks_per_s": round(st["blocks_allocated"] / mc_s, 1),
                    "rot_err_deg": {"dead_reckoning": summ(e_dead[:, 0]), "tracked": summ(e_trk[:, 0])},
                    "trans_err_mm": {"dead_reckoning": summ(e_dead[:, 1]), "tracked": summ(e_trk[:, 1])}})
        # per-stage device time (eager launches with events, same sequence)
        vol.reset()
        vol.set_profiling(True)
        vol.bind_tracking(d_depth, vh.DEPTH_U16, 0.001, nf, d_inc, d_out, d_stats, gn_iters, stride, zmax, mode,
                          **motion)
        vol.integrate_next(nf)
        vol.sync()
        p = vol.stats()
        vol.set_profiling(False)
        n = max(1, p["profiled_frames"])
        res["device_us_per_frame"] = {"track": round(1e3 * p["track_ms"] / n, 2),
                                      "alloc": round(1e3 * p["alloc_ms"] / n, 2),
                                      "integrate": round(1e3 * p["integrate_ms"] / n, 2)}
        log(f"bench: tracked: {res['fps']} frames/s, {st['blocks_allocated']} blocks; mean error dead reckoning "
            f"{res['rot_err_deg']['dead_reckoning']['mean']} deg {res['trans_err_mm']['dead_reckoning']['mean']} mm, "
            f"tracked {res['rot_err_deg']['tracked']['mean']} deg {res['trans_err_mm']['tracked']['mean']} mm")
    except vh.VhError as e:
        res["error"] = str(e)
        log(f"bench: tracked: {e}")
    vol.close()
    return res


def run_pass(vol, vh, seq, nf, mc_every, d_xyz, cap, final_mc=True, d_rgb=None, d_vrgb=None):
    """One pass over the bound sequence from an empty volume, extracting the mesh every
    mc_every frames and (if final_mc) at the end; returns (last triangle count, extraction
    wall times). d_rgb: the sequence's RGB frames (colour volumes); d_vrgb: vertex colours."""
    d_depth, d_pose = seq
    vol.reset()
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, nf)
    if d_rgb is not None:
        vol.bind_sequence_rgb(d_rgb)
    mcs = []
    ntri = 0
    f = 0
    while f < nf:
        n = min(mc_every, nf - f) if mc_every else nf - f
        vol.integrate_next(n)
        f += n
        if mc_every or (final_mc and f == nf):
            t0 = time.perf_counter()
            ntri = vol.extract_mesh_device(d_xyz, None, cap, d_vrgb)  # waits for the frames and the extraction
            mcs.append((t0, time.perf_counter()))
    vol.sync()
    return ntri, mcs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="room_medium_5mm")
    ap.add_argument("--frames", type=int, default=0, help="override the sequence length")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f16", action="store_true")
    ap.add_argument("--no-color", action="store_true", help="skip the colour variants")
    ap.add_argument("--seq-batch", type=int, default=0,
                    help="frames per integration launch on the bound sequence (1..8; 0 = the library's 8)")
    ap.add_argument("--no-gn", action="store_true", help="skip the pose-refinement block")
    ap.add_argument("--no-tracked", action="store_true", help="skip the tracked floor_large_4mm block")
    ap.add_argument("--tracked-frames", type=int, default=0, help="override the tracked sequence length")
    ap.add_argument("--tracked-only", action="store_true",
                    help="run only the tracked floor_large_4mm block (prints its JSON object)")
    ap.add_argument("--counters", action="store_true",
                    help="short run for hardware-counter passes: one FP32 integration pass over the "
                         "sequence plus a 1 GiB device copy that calibrates FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--counters-fused", action="store_true",
                    help="with --counters: the captured-graph batches (fused integration + allocation launches) "
                         "instead of the profiling mode's separate launches")
    ap.add_argument("--counters-mc", action="store_true",
                    help="with --counters: one marching-cubes extraction after the integration pass")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: libvh does not shard a volume across GPUs "
              "(sharding would be by space, one owner GPU per block)", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__

    vh, wl = __graft_entry__.package()
    vh.load()
    if args.tracked_only:
        print(json.dumps({"tracked": tracked_block(vh, wl, args.tracked_frames)}), flush=True)
        return
    w = dict(wl.WORKLOADS[args.workload])
    if args.seq_batch:
        w["seq_batch"] = args.seq_batch
    nf = args.frames or w["frames"]
    mc_every = w.get("mc_every", 0)
    log(f"bench: workload {args.workload}, {nf} frames, steps {args.steps}, warmup {args.warmup}")

    scene = vh.Scene(w["scene"], wl.SEED)
    t = time.perf_counter()
    depth, poses = scene.frames(0, nf)
    log(f"bench: generated {nf} frames in {time.perf_counter() - t:.1f} s")
    d_depth = torch.from_numpy(depth.view(np.int16)).cuda()
    d_pose = torch.from_numpy(poses).cuda()
    cap = 48 << 20  # triangles (1.7 GB of float xyz)
    d_xyz = torch.empty(cap * 9, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    if args.counters:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        vol = vh.Volume(vh.config_from_workload(w))
        # allocation and integration as separate launches per batch, or the fused graph batches
        vol.set_profiling(2 if args.counters_fused else True)
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=args.counters_mc)
        st = vol.stats()
        vol.close()
        launches = st["fused_launches"] if args.counters_fused else st["profiled_launches"]
        log(f"bench: counters run: {nf} frames in {launches} {'fused' if args.counters_fused else 'integration'} "
            f"launches, {st['blocks_integrated']} blocks integrated, copy calibration 1 GiB read + 1 GiB written")
        print(json.dumps({"counters_run": True, "frames": nf, "touched_total": st["touched_total"],
                          "blocks_integrated": st["blocks_integrated"], "launches": launches,
                          "fused": bool(args.counters_fused), "copy_bytes": 1 << 30}), flush=True)
        return

    results = {}
    # host frames from page-locked memory: copy ring + allocation + integration per frame.
    # Measured first: a large host allocation made just before (the colour frames below) has
    # been seen to halve this rate for the next seconds.
    vol = vh.Volume(vh.config_from_workload(w))
    pinned = torch.from_numpy(depth.view(np.int16)).pin_memory()
    for k in range(min(20, nf)):
        vol.integrate_async(pinned[k], poses[k])
    vol.sync()
    vol.reset()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for k in range(nf):
        vol.integrate_async(pinned[k], poses[k])
    vol.sync()
    results["pcie"] = nf / (time.perf_counter() - t3)
    vol.close()
    del pinned
    log(f"bench: host input: {results['pcie']:.1f} frames/s")

    variants = [("f32", {}), ("f32v", dict(carve=1))]
    if not args.no_f16:
        variants += [("f16r", dict(storage=2, weight_cap=255)), ("f16", dict(storage=1, weight_cap=255))]
    d_rgb = d_vrgb = None
    if not args.no_color:
        variants += [("f32c", dict(color=1))]
        if not args.no_f16:
            variants += [("f16rc", dict(storage=2, weight_cap=255, color=1))]
        t = time.perf_counter()
        rgb = np.stack([scene.rgb(k, poses[k]) for k in range(nf)])
        log(f"bench: rendered {nf} colour frames in {time.perf_counter() - t:.1f} s")
        d_rgb = torch.from_numpy(rgb).cuda()
        del rgb
        d_vrgb = torch.empty(cap * 9, dtype=torch.uint8, device="cuda")
    for name, over in variants:
        col = "color" in over
        cr = dict(d_rgb=d_rgb, d_vrgb=d_vrgb) if col else {}
        vol = vh.Volume(vh.config_from_workload(w, **over))
        for _ in range(args.warmup):
            run_pass(vol, vh, (d_depth, d_pose), nf, mc_every, d_xyz, cap, **cr)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ntri, mcs = run_pass(vol, vh, (d_depth, d_pose), nf, mc_every, d_xyz, cap, **cr)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = vol.stats()
        if ntri > cap:
            log(f"bench: WARNING mesh of {ntri} triangles exceeds the buffer ({cap})")
        # integration only (no extraction), same sequence
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False, **cr)
        torch.cuda.synchronize()
        dt_int = time.perf_counter() - t1
        # final extraction alone, timed with the frames already integrated
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ntri = vol.extract_mesh_device(d_xyz, None, cap, d_vrgb if col else None)
        mc_s = (time.perf_counter() - t2) / reps
        ix = None
        if name == "f32":  # the indexed mesh of the same volume (device arrays sized by a first call)
            nv, nfc, _ = vol.extract_mesh_indexed_device(None, 0, None, 0)
            d_iv = torch.empty((max(nv, 1), 3), dtype=torch.float32, device="cuda")
            d_if = torch.empty((max(nfc, 1), 3), dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for _ in range(reps):
                nv, nfc, rc = vol.extract_mesh_indexed_device(d_iv, nv, d_if, nfc)
            ix = dict(ms=1e3 * (time.perf_counter() - t2) / reps, vertices=nv, faces=nfc, ok=rc == vh.VH_OK)
            del d_iv, d_if
        # kernel timing with HIP events (eager launches on the volume's stream)
        vol.set_profiling(True)
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False, **cr)
        pst = vol.stats()
        vol.set_profiling(False)
        fst = None
        if name == "f32":  # the fused batch launches of the captured graphs, HIP events around each
            vol.set_profiling(2)
            run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False, **cr)
            fst = vol.stats()
            vol.set_profiling(False)
        stream = None
        if name == "f32":  # the whole volume to the host store and back (after the profiled pass)
            t2 = time.perf_counter()
            n_out = vol.stream_out((0.0, 0.0, 0.0), 0.0)
            t_out = time.perf_counter() - t2
            t2 = time.perf_counter()
            n_in = vol.stream_in()
            t_in = time.perf_counter() - t2
            stream = dict(blocks=n_out, back=n_in, out_s=t_out, in_s=t_in)
        if name == "f16" and not args.no_gn:  # the volume now holds the whole sequence
            results["gn"] = gn_block(vh, wl, w, vol, nf)
        results[name] = dict(fps=args.steps * nf / dt, ms_per_step=1e3 * dt / args.steps, fps_integrate=nf / dt_int,
                             blocks=st["blocks_allocated"], tris=ntri, mc_ms=1e3 * mc_s,
                             mc_blocks_per_s=st["blocks_allocated"] / mc_s, prof=pst, fused=fst, indexed=ix,
                             stream=stream,
                             vox_bytes=(8 if over.get("storage", 0) == 0 else 4) + (4 if col else 0))
        log(f"bench: {name}: {results[name]['fps']:.1f} frames/s (with MC), {results[name]['fps_integrate']:.1f} "
            f"integrate-only, {st['blocks_allocated']} blocks, {ntri} triangles, MC {1e3 * mc_s:.2f} ms")
        vol.close()

    r = results["f32"]
    p = r["prof"]
    W, H = w["width"], w["height"]
    # per integration launch (one batch): its frames' depth images + each listed block's voxels
    # read and written once + its 16-B list entry and slot value
    bytes_total = p["profiled_frames"] * W * H * 2 + p["blocks_integrated"] * (512 * r["vox_bytes"] * 2 + 16)
    t_int = p["integrate_ms"] / 1e3
    launches = p["profiled_launches"]
    achieved = bytes_total / t_int / 1e9
    traffic, traffic_src = None, None
    prof = os.path.join(ROOT, "profiles", "k_integrate_room_medium_5mm.json")
    pj = json.load(open(prof)) if os.path.exists(prof) else {}
    # the summary must describe the same launches (workload, frames per launch) to be used
    if (args.workload == "room_medium_5mm" and "blocks_per_dispatch" in pj
            and abs(pj["frames_per_dispatch"] - p["profiled_frames"] / launches) < 1e-6):
        traffic = int(pj["traffic_bytes_per_dispatch"])
        traffic_src = (f"rocprofv3 FETCH_SIZE x{pj['fetch_scale']:.3f} + WRITE_SIZE x{pj['write_scale']:.3f} "
                       f"(1 GiB copy calibration), {pj['workload']}, "
                       f"{pj['blocks_per_dispatch']:.0f} blocks and {pj['frames_per_dispatch']:.0f} frames per launch")
    valu = None
    if traffic is not None and "valu_busy" in pj:
        valu = {"busy": round(pj["valu_busy"], 3),
                "ta_busy": round(pj["ta_busy"], 3) if "ta_busy" in pj else None,
                "ta_stalled_by_l1": round(pj["ta_stalled_by_tc"], 3) if "ta_stalled_by_tc" in pj else None,
                "source": "rocprofv3 SQ_ACTIVE_INST_VALU, TA_BUSY_avr, TA_ADDR_STALLED_BY_TC_CYCLES / GRBM_GUI_ACTIVE "
                          "of the same launches: each block's depth window is loaded into LDS once per frame, so "
                          "the texture path is no longer stalled; the launch is bound by vector instruction issue "
                          "(projection, correctly rounded reciprocal and update per voxel and frame), not by HBM"}
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src, valu=valu,
                    kernel="k_integrate<F32,u16>", avg_launch_us=round(1e6 * t_int / launches, 2),
                    bytes_per_launch=int(bytes_total / launches),
                    frames_per_launch=round(p["profiled_frames"] / launches, 3),
                    blocks_per_launch=round(p["blocks_integrated"] / launches, 1),
                    touched_blocks_per_frame=round(p["touched_total"] / p["profiled_frames"], 1),
                    integrate_us_per_frame=round(1e6 * t_int / p["profiled_frames"], 2),
                    alloc_avg_launch_us=round(1e6 * p["alloc_ms"] / 1e3 / launches, 2))

    # the fused launch (integration of batch b + allocation of batch b+1): algorithmic bytes =
    # the integration's (as above) + the allocation's, one depth image per frame plus 16 B per
    # touched block (hash probe and list entry); time from HIP events around each graph launch
    roofline_fused = None
    fz = r.get("fused")
    if fz and fz["fused_launches"]:
        nb = p["blocks_integrated"] / launches
        fpl = p["profiled_frames"] / launches
        tb = p["touched_total"] / p["profiled_frames"]
        fused_bytes = nb * (512 * r["vox_bytes"] * 2 + 16) + fpl * W * H * 2 + fpl * (W * H * 2 + 16 * tb)
        t_f = fz["fused_ms"] / 1e3 / fz["fused_launches"]
        a_f = fused_bytes / t_f / 1e9
        fj_path = os.path.join(ROOT, "profiles", "k_integrate_alloc_room_medium_5mm.json")
        fj = json.load(open(fj_path)) if os.path.exists(fj_path) else {}
        f_traffic, f_src = None, None
        if args.workload == "room_medium_5mm" and "traffic_bytes_per_dispatch" in fj:
            f_traffic = int(fj["traffic_bytes_per_dispatch"])
            f_src = (f"rocprofv3 FETCH_SIZE x{fj['fetch_scale']:.3f} + WRITE_SIZE x{fj['write_scale']:.3f} "
                     f"(1 GiB copy calibration), {fj['workload']}, {fj['dispatches']} fused launches")
        f_valu = None
        if f_traffic is not None and "valu_busy" in fj:
            f_valu = {"busy": round(fj["valu_busy"], 3),
                      "source": "rocprofv3 SQ_ACTIVE_INST_VALU / GRBM_GUI_ACTIVE of the same launches: the fused "
                                "launch is bound by vector instruction issue (integration + band walks)"}
        roofline_fused = dict(bound="hbm", achieved=round(a_f, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=round(a_f / HBM_PEAK_GBS, 4), traffic=f_traffic, traffic_source=f_src, valu=f_valu,
                              kernel="k_integrate_alloc<F32,u16>", avg_launch_us=round(1e6 * t_f, 2),
                              timing="HIP events around each captured-graph batch launch on the volume's stream",
                              bytes_per_launch=int(fused_bytes), launches=fz["fused_launches"])
    out = {"metric": METRIC, "value": round(r["fps"], 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": args.workload, "frames_per_step": nf, "width": W, "height": H,
                      "voxel_m": w["vox_size"], "trunc_m": w["trunc"], "log2_buckets": w["log2_buckets"],
                      "pool_blocks": w["pool_blocks"], "block": 8, "mc_every": mc_every,
                      "depth": "uint16 mm, device-resident",
                      "per_launch": f"one captured-graph launch per batch of {w.get('seq_batch', 8)} frames: integration of batch b "
                                    "(each block read and written once, updated by each of its frames in "
                                    "order) and allocation of batch b+1"},
           "integrate_only_fps": round(r["fps_integrate"], 2), "mc_blocks": r["blocks"],
           "mc_triangles": r["tris"], "mc_ms": round(r["mc_ms"], 3), "mc_blocks_per_s": round(r["mc_blocks_per_s"], 1),
           "roofline": roofline, "roofline_fused": roofline_fused}
    if r.get("indexed"):
        ix = r["indexed"]
        out["mesh_indexed"] = {"ms": round(ix["ms"], 3), "vertices": ix["vertices"], "faces": ix["faces"],
                               "faces_per_s": round(ix["faces"] / (ix["ms"] / 1e3), 1), "complete": ix["ok"],
                               "vs_soup_ms": round(ix["ms"] / r["mc_ms"], 3),
                               "note": "vh_extract_mesh_indexed_device on the final volume: used-edge bits per "
                                       "block (mark pass), vertex ids from per-block popcount offsets, faces and "
                                       "vertices in one write pass; one vertex per crossed lattice edge"}
    if r.get("stream"):
        sm = r["stream"]
        gb = sm["blocks"] * 512 * r["vox_bytes"] / 1e9
        out["streaming"] = {"blocks": sm["blocks"], "gbytes": round(gb, 3),
                            "out_ms": round(1e3 * sm["out_s"], 1), "in_ms": round(1e3 * sm["in_s"], 1),
                            "out_gbs": round(gb / sm["out_s"], 2), "in_gbs": round(gb / sm["in_s"], 2),
                            "complete": sm["back"] == sm["blocks"],
                            "note": "vh_stream_out of every block of the final volume to the host store "
                                    "(radius 0), then vh_stream_in of all of them: voxel words over PCIe, "
                                    "pool compaction, hash rebuild"}
    labels = {"f16": ("fp16", "binary16 tsdf + uint16 weight (cap 255), 4 B/voxel"),
              "f16r": ("f16r", "F16R: binary16 tsdf + uint8 weight (cap 255) + int8 rounding residual "
                               "(19 significant bits), 4 B/voxel")}
    if "pcie" in results:
        out["host_input_fps"] = {"value": round(results.pop("pcie"), 2), "unit": "frames/s",
                                 "note": "the same 500 frames from page-locked host memory: vh_integrate_async per frame "
                                         "(uint16 depth copied over PCIe into a 4-slot device ring on a second stream, "
                                         "then allocation + integration), no marching cubes"}
    if "gn" in results:
        out["gn"] = results.pop("gn")
    for name, (key, desc) in labels.items():
        if name not in results:
            continue
        h = results[name]
        ph = h["prof"]
        bh = ph["profiled_frames"] * W * H * 2 + ph["blocks_integrated"] * (512 * 4 * 2 + 16)
        ah = bh / (ph["integrate_ms"] / 1e3) / 1e9
        out[key] = {"storage": desc, "value": round(h["fps"], 2), "integrate_only_fps": round(h["fps_integrate"], 2),
                    "speedup_vs_f32": round(h["fps"] / r["fps"], 3),
                    "integrate_speedup_vs_f32": round(h["fps_integrate"] / r["fps_integrate"], 3),
                    "mc_ms": round(h["mc_ms"], 3), "mc_blocks_per_s": round(h["mc_blocks_per_s"], 1),
                    "integrate_kernel_gbs": round(ah, 1),
                    "integrate_avg_launch_us": round(1e6 * ph["integrate_ms"] / 1e3 / ph["profiled_launches"], 2)}

    if "f32v" in results:  # free-space carving (vh_config.carve = 1) against band-only integration
        h = results["f32v"]
        ph = h["prof"]
        out["carve"] = {"value": round(h["fps"], 2), "integrate_only_fps": round(h["fps_integrate"], 2),
                        "vs_band_only": round(h["fps"] / r["fps"], 3),
                        "integrate_vs_band_only": round(h["fps_integrate"] / r["fps_integrate"], 3),
                        "carve_us_per_batch": round(1e6 * ph["carve_ms"] / 1e3 / ph["profiled_launches"], 2),
                        "integrate_avg_launch_us": round(1e6 * ph["integrate_ms"] / 1e3 / ph["profiled_launches"], 2),
                        "blocks_integrated_per_batch": round(ph["blocks_integrated"] / ph["profiled_launches"], 1),
                        "carved_block_frames": ph["carved_total"], "mc_ms": round(h["mc_ms"], 3),
                        "note": "FP32 with carving: after each batch's allocation k_carve lists every resident block "
                                "a frame may update, so voxels in front of the surface move toward +1"}

    colour_labels = {"f32c": ("color", "f32", "fp32 tsdf/weight (8 B) + colour word (4 B) per voxel"),
                     "f16rc": ("color_fp16", "f16r", "F16R tsdf/weight (4 B) + colour word (4 B) per voxel")}
    for name, (key, base, desc) in colour_labels.items():
        if name not in results:
            continue
        h = results[name]
        ph = h["prof"]
        # per touched block: voxel and colour words read and written; per frame: depth and RGB
        bh = ph["profiled_frames"] * W * H * (2 + 4) + ph["blocks_integrated"] * (512 * h["vox_bytes"] * 2 + 16)
        ah = bh / (ph["integrate_ms"] / 1e3) / 1e9
        b = results[base]
        out[key] = {"storage": desc, "value": round(h["fps"], 2), "integrate_only_fps": round(h["fps_integrate"], 2),
                    "vs_depth_only": round(h["fps"] / b["fps"], 3),
                    "integrate_vs_depth_only": round(h["fps_integrate"] / b["fps_integrate"], 3),
                    "mc_ms": round(h["mc_ms"], 3), "mc_vs_depth_only": round(h["mc_ms"] / b["mc_ms"], 3),
                    "integrate_kernel_gbs": round(ah, 1),
                    "integrate_avg_launch_us": round(1e6 * ph["integrate_ms"] / 1e3 / ph["profiled_launches"], 2),
                    "rgb": "synthetic scene colours (vh_scene_render_rgb), device-resident"}

    if not args.no_tracked:
        del d_xyz, d_depth, d_pose, d_rgb, d_vrgb
        torch.cuda.empty_cache()
        out["tracked"] = tracked_block(vh, wl, args.tracked_frames)
    if not args.no_cpu:
        out["cpu_baseline"], errs = cpu_baseline(vh, w, depth, poses, args.cpu_seconds, args.no_f16, wl)
        for key, err in errs.items():
            out.setdefault(key, {}).update(err)
    # the metric's two targets for the 4-byte formats: 1.5x the FP32 frames/s, TSDF error <= 1e-3
    for key in ("fp16", "f16r"):
        if key in out and "speedup_vs_f32" in out[key]:
            t = {"speedup_vs_f32_ge_1.5": out[key]["speedup_vs_f32"] >= 1.5}
            if "tsdf_err_max" in out[key]:
                t["tsdf_err_max_le_1e-3"] = out[key]["tsdf_err_max"] <= 1e-3
            out[key]["targets_met"] = t
    print(json.dumps(out), flush=True)


def jacobian_vs_fd(vh, wl, w, vol, ref, frame, h=1e-4, max_points=20000):
    """GPU pose Jacobians (left se(3) perturbation, vh_query_sdf_backward_device) at the points
    of a held-out frame under a perturbed pose, against central differences (step h) of the
    oracle's FP64 trilinear query on the same voxel values, over points whose stencil stays in
    one lattice cell; returns the relative error and the points used."""
    import torch

    scene = vh.Scene(w["scene"], wl.SEED)
    d, T = scene.frames(frame, 1)
    rng = np.random.default_rng(wl.SEED + 2)
    xi = np.concatenate([rng.normal(0, 0.03, 3), rng.normal(0, np.deg2rad(2.0), 3)])
    T0 = (T[0].astype(np.float64) @ se3_exp(0.25 * xi)).astype(np.float32)
    pts = backproject(d[0], w, zmax=3.0)
    n = len(pts)
    d_pts = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    d_r = torch.zeros(n, dtype=torch.float32, device="cuda")
    d_valid = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_J = torch.zeros((n, 6), dtype=torch.float32, device="cuda")
    vol.backward_device(d_pts, n, T0, None, d_r, d_valid, d_J)
    torch.cuda.synchronize()
    valid, J = d_valid.cpu().numpy().astype(bool), d_J.cpu().numpy().astype(np.float64)
    q = pts.astype(np.float64) @ T0[:3, :3].astype(np.float64).T + T0[:3, 3].astype(np.float64)
    q, J = q[valid][:max_points], J[valid][:max_points]
    vox = float(np.float32(w["vox_size"]))
    cell = np.floor(q / vox)
    ok = np.ones(len(q), bool)
    fd = np.zeros((len(q), 6))
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Ep, Em = se3_exp(e), se3_exp(-e)
        qp, qm = q @ Ep[:3, :3].T + Ep[:3, 3], q @ Em[:3, :3].T + Em[:3, 3]
        ok &= np.all(np.floor(qp / vox) == cell, 1) & np.all(np.floor(qm / vox) == cell, 1)
        sp, _, vp = ref.query64(qp)
        sm, _, vm = ref.query64(qm)
        ok &= vp & vm
        fd[:, k] = (sp - sm) / (2 * h)
    rel = float(np.linalg.norm(J[ok] - fd[ok]) / np.linalg.norm(fd[ok]))
    return {"rel_err": rel, "points": int(ok.sum()), "frame": frame, "h": h,
            "reference": "central differences of the oracle's FP64 trilinear query on the same voxel "
                         "values (left se(3) perturbation), points of a held-out frame at a pose "
                         "perturbed by 1/4 of the gn error, stencils within one cell"}


def cpu_baseline(vh, w, depth, poses, budget_s, no_f16, wl=None):
    """The oracle (same algorithm: FP32 twin + FP64 shadow, OpenMP) on the first frames of the
    same sequence, as many as fit the time budget; then the GPU FP16 volume after the same
    frames against the FP64 shadow."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    cores = oracle.load().vho_max_threads()
    twin = oracle.Oracle(w, shadow=False)
    t0 = time.perf_counter()
    n = 0
    while n < len(depth) and (time.perf_counter() - t0 < budget_s or n < 5):
        twin.integrate(depth[n], poses[n])
        n += 1
    dt = time.perf_counter() - t0
    del twin
    # the same frames with the FP64 shadow as well (the reference the errors below use)
    ref = oracle.Oracle(w)
    t0 = time.perf_counter()
    for k in range(n):
        ref.integrate(depth[k], poses[k])
    dt2 = time.perf_counter() - t0
    base = {"value": round(n / dt, 3), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"first {n} frames of the workload, the FP32 twin alone (the GPU path's arithmetic), "
                      f"OpenMP {cores} threads (hardware_concurrency {os.cpu_count()}), {dt:.1f} s",
            "with_fp64_shadow": {"value": round(n / dt2, 3), "unit": "frames/s",
                                 "sample": f"the same {n} frames, FP32 twin + FP64 shadow, {dt2:.1f} s"}}
    errs = {}
    rk, rs, rw, r64, rw64 = ref.export()
    t = time.perf_counter()
    m64 = ref.mc64()  # the FP64 reference mesh of the same frames
    log(f"bench: FP64 reference mesh of {n} frames: {len(m64[0])} triangles in {time.perf_counter() - t:.1f} s")
    storages = [("fp32", {})]
    if not no_f16:
        storages += [("f16r", dict(storage=2, weight_cap=255)), ("fp16", dict(storage=1, weight_cap=255))]
    for key, over in storages:
        vol = vh.Volume(vh.config_from_workload(w, **over))
        for k in range(n):
            vol.integrate(depth[k], poses[k])
        keys, s, wt = vol.export_blocks()
        keys, s, wt = oracle.sort_blocks(keys, s, wt)
        if np.array_equal(keys, rk):
            m = wt > 0
            e = np.abs(s[m].astype(np.float64) - r64[m])
            errs[key] = {"tsdf_err_max": float(e.max()), "tsdf_err_rms": float(np.sqrt((e ** 2).mean())),
                         "err_frames": n, "err_reference": "oracle FP64 shadow, all observed voxels"}
        else:
            errs[key] = {"tsdf_err": "block sets differ from the oracle"}
        t = time.perf_counter()
        xyz, cubes = vol.extract_mesh()
        h = oracle.hausdorff(xyz, cubes, *m64)
        errs[key].update({"hausdorff_m": h["hausdorff"], "hausdorff_q0.999_m": h["q0.999"],
                          "hausdorff_q0.99999_m": h["q0.99999"], "mesh_to_fp64_mean_m": h["a_to_b"]["mean"],
                          "fp64_to_mesh_mean_m": h["b_to_a"]["mean"],
                          "hausdorff_unmatched_samples": h["a_to_b"]["far"] + h["b_to_a"]["far"],
                          "hausdorff_reference": "GPU mesh vs marching cubes over the oracle's FP64 shadow; "
                                                 "samples = vertices + centroids of both meshes, exact "
                                                 "point-to-triangle distances within the 27 surrounding cubes"})
        log(f"bench: {key}: Hausdorff to the FP64 mesh {1e3 * h['hausdorff']:.4f} mm "
            f"({time.perf_counter() - t:.1f} s)")
        if key == "fp32" and wl is not None and np.array_equal(keys, rk):
            g = jacobian_vs_fd(vh, wl, w, vol, ref, n + 17)
            errs[key]["pose_jacobian_vs_fd"] = g
            log(f"bench: fp32 pose Jacobian vs FD: relative error {g['rel_err']:.2e} over {g['points']} points")
        vol.close()
    # carving: the GPU's FP32 carving volume after the first frames against the oracle's FP64
    # shadow in carving mode (which tests every allocated block each frame)
    nc = min(n, 150)
    t = time.perf_counter()
    cref = oracle.Oracle(w, carve=1)
    for k in range(nc):
        cref.integrate(depth[k], poses[k])
    vol = vh.Volume(vh.config_from_workload(w, carve=1))
    for k in range(nc):
        vol.integrate(depth[k], poses[k])
    keys, s, wt = oracle.sort_blocks(*vol.export_blocks())
    rk, rs, rw, r64, _ = cref.export()
    c = {"err_frames": nc}
    if np.array_equal(keys, rk):
        m = wt > 0
        e = np.abs(s[m].astype(np.float64) - r64[m])
        c.update(tsdf_err_max=float(e.max()), tsdf_err_rms=float(np.sqrt((e ** 2).mean())),
                 bit_exact_vs_fp32_twin=bool(np.array_equal(s.view(np.uint32), rs.view(np.uint32))))
        xyz, cubes = vol.extract_mesh()
        h = oracle.hausdorff(xyz, cubes, *cref.mc64())
        c.update({"hausdorff_m": h["hausdorff"], "hausdorff_q0.999_m": h["q0.999"],
                  "hausdorff_q0.99999_m": h["q0.99999"]})
    else:
        c["tsdf_err"] = "block sets differ from the oracle"
    vol.close()
    errs["carve"] = c
    log(f"bench: carving, first {nc} frames: {c} ({time.perf_counter() - t:.1f} s)")
    return base, errs


if __name__ == "__main__":
    main()
