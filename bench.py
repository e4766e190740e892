#!/usr/bin/env python3
"""Benchmark: integrated frames/s and marching-cubes blocks/s on one MI355X.

Workload (default room_medium_5mm): 500 synthetic 640x480 uint16 depth frames of an
8x6x3 m room with 12 primitives, 5 mm voxels, 2 cm truncation, 2^21 buckets, 1.5M-block
pool, FP32 voxels; per-frame work is one captured HIP graph (allocation + integration);
marching cubes runs every 100 frames (5 extractions per pass, the last one final).
One step = one pass over the whole sequence from an empty volume; the depth frames and
poses are resident in HBM before timing starts.

Reported beside the headline FP32 figure: the same pass with binary16 TSDF + uint16 weight
storage, the FP16/FP32 speedup and the FP16 error against the oracle's FP64 shadow, the
integration kernel's achieved HBM bandwidth (algorithmic bytes / kernel time, measured with
HIP events on the volume's stream), and the oracle's CPU throughput on a bounded sample.
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


def run_pass(vol, vh, seq, nf, mc_every, d_xyz, cap, final_mc=True):
    """One pass over the bound sequence from an empty volume, extracting the mesh every
    mc_every frames and (if final_mc) at the end; returns (last triangle count, extraction
    wall times)."""
    d_depth, d_pose = seq
    vol.reset()
    vol.bind_sequence(d_depth, vh.DEPTH_U16, 0.001, d_pose, nf)
    mcs = []
    ntri = 0
    f = 0
    while f < nf:
        n = min(mc_every, nf - f) if mc_every else nf - f
        vol.integrate_next(n)
        f += n
        if mc_every or (final_mc and f == nf):
            t0 = time.perf_counter()
            ntri = vol.extract_mesh_device(d_xyz, None, cap)  # waits for the frames and the extraction
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
    ap.add_argument("--counters", action="store_true",
                    help="short run for hardware-counter passes: one FP32 integration pass over the "
                         "sequence plus a 1 GiB device copy that calibrates FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the volume is not sharded across GPUs in this build "
              "(it would shard by space, one owner GPU per block)", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__

    vh, wl = __graft_entry__.package()
    vh.load()
    w = dict(wl.WORKLOADS[args.workload])
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
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False)
        st = vol.stats()
        vol.close()
        log(f"bench: counters run: {nf} frames, {st['touched_total']} touched blocks in total, "
            f"copy calibration 1 GiB read + 1 GiB written")
        print(json.dumps({"counters_run": True, "frames": nf, "touched_total": st["touched_total"],
                          "copy_bytes": 1 << 30}), flush=True)
        return

    results = {}
    variants = [("f32", {})]
    if not args.no_f16:
        variants += [("f16r", dict(storage=2, weight_cap=255)), ("f16", dict(storage=1, weight_cap=255))]
    for name, over in variants:
        vol = vh.Volume(vh.config_from_workload(w, **over))
        for _ in range(args.warmup):
            run_pass(vol, vh, (d_depth, d_pose), nf, mc_every, d_xyz, cap)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ntri, mcs = run_pass(vol, vh, (d_depth, d_pose), nf, mc_every, d_xyz, cap)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = vol.stats()
        if ntri > cap:
            log(f"bench: WARNING mesh of {ntri} triangles exceeds the buffer ({cap})")
        # integration only (no extraction), same sequence
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False)
        torch.cuda.synchronize()
        dt_int = time.perf_counter() - t1
        # final extraction alone, timed with the frames already integrated
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ntri = vol.extract_mesh_device(d_xyz, None, cap)
        mc_s = (time.perf_counter() - t2) / reps
        # kernel timing with HIP events (eager launches on the volume's stream)
        vol.set_profiling(True)
        run_pass(vol, vh, (d_depth, d_pose), nf, 0, d_xyz, cap, final_mc=False)
        pst = vol.stats()
        vol.set_profiling(False)
        results[name] = dict(fps=args.steps * nf / dt, ms_per_step=1e3 * dt / args.steps, fps_integrate=nf / dt_int,
                             blocks=st["blocks_allocated"], tris=ntri, mc_ms=1e3 * mc_s,
                             mc_blocks_per_s=st["blocks_allocated"] / mc_s, prof=pst,
                             vox_bytes=8 if name == "f32" else 4)
        log(f"bench: {name}: {results[name]['fps']:.1f} frames/s (with MC), {results[name]['fps_integrate']:.1f} "
            f"integrate-only, {st['blocks_allocated']} blocks, {ntri} triangles, MC {1e3 * mc_s:.2f} ms")
        vol.close()

    r = results["f32"]
    p = r["prof"]
    W, H = w["width"], w["height"]
    bytes_total = p["profiled_frames"] * W * H * 2 + p["touched_total"] * (512 * r["vox_bytes"] * 2 + 16)
    t_int = p["integrate_ms"] / 1e3
    achieved = bytes_total / t_int / 1e9
    traffic, traffic_src = None, None
    prof = os.path.join(ROOT, "profiles", "k_integrate_room_medium_5mm.json")
    if args.workload == "room_medium_5mm" and os.path.exists(prof):
        pj = json.load(open(prof))
        traffic = int(pj["traffic_bytes_per_dispatch"])
        traffic_src = (f"rocprofv3 FETCH_SIZE x{pj['fetch_scale']:.3f} + WRITE_SIZE x{pj['write_scale']:.3f} "
                       f"(1 GiB copy calibration), {pj['workload']}, "
                       f"{pj['touched_blocks_per_dispatch']:.0f} touched blocks per launch")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                    kernel="k_integrate<F32,u16>", avg_launch_us=round(1e6 * t_int / p["profiled_frames"], 2),
                    bytes_per_launch=int(bytes_total / p["profiled_frames"]),
                    touched_blocks_per_frame=round(p["touched_total"] / p["profiled_frames"], 1),
                    alloc_avg_launch_us=round(1e6 * p["alloc_ms"] / 1e3 / p["profiled_frames"], 2))

    out = {"metric": METRIC, "value": round(r["fps"], 2), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(r["ms_per_step"], 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": args.workload, "frames_per_step": nf, "width": W, "height": H,
                      "voxel_m": w["vox_size"], "trunc_m": w["trunc"], "log2_buckets": w["log2_buckets"],
                      "pool_blocks": w["pool_blocks"], "block": 8, "mc_every": mc_every,
                      "depth": "uint16 mm, device-resident", "per_frame": "captured HIP graph"},
           "integrate_only_fps": round(r["fps_integrate"], 2), "mc_blocks": r["blocks"],
           "mc_triangles": r["tris"], "mc_ms": round(r["mc_ms"], 3), "mc_blocks_per_s": round(r["mc_blocks_per_s"], 1),
           "roofline": roofline}
    labels = {"f16r": ("fp16", "binary16 tsdf + uint8 weight (cap 255) + int8 rounding residual, 4 B/voxel"),
              "f16": ("fp16_plain", "binary16 tsdf + uint16 weight (cap 255), 4 B/voxel")}
    for name, (key, desc) in labels.items():
        if name not in results:
            continue
        h = results[name]
        ph = h["prof"]
        bh = ph["profiled_frames"] * W * H * 2 + ph["touched_total"] * (512 * 4 * 2 + 16)
        ah = bh / (ph["integrate_ms"] / 1e3) / 1e9
        out[key] = {"storage": desc, "value": round(h["fps"], 2), "integrate_only_fps": round(h["fps_integrate"], 2),
                    "speedup_vs_f32": round(h["fps"] / r["fps"], 3),
                    "integrate_speedup_vs_f32": round(h["fps_integrate"] / r["fps_integrate"], 3),
                    "mc_ms": round(h["mc_ms"], 3), "mc_blocks_per_s": round(h["mc_blocks_per_s"], 1),
                    "integrate_kernel_gbs": round(ah, 1),
                    "integrate_avg_launch_us": round(1e6 * ph["integrate_ms"] / 1e3 / ph["profiled_frames"], 2)}

    if not args.no_cpu:
        out["cpu_baseline"], errs = cpu_baseline(vh, w, depth, poses, args.cpu_seconds, args.no_f16)
        for key, err in errs.items():
            out.setdefault(key, {}).update(err)
    print(json.dumps(out), flush=True)


def cpu_baseline(vh, w, depth, poses, budget_s, no_f16):
    """The oracle (same algorithm: FP32 twin + FP64 shadow, OpenMP) on the first frames of the
    same sequence, as many as fit the time budget; then the GPU FP16 volume after the same
    frames against the FP64 shadow."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    ref = oracle.Oracle(w)
    cores = oracle.load().vho_max_threads()
    t0 = time.perf_counter()
    n = 0
    while n < len(depth) and (time.perf_counter() - t0 < budget_s or n < 5):
        ref.integrate(depth[n], poses[n])
        n += 1
    dt = time.perf_counter() - t0
    base = {"value": round(n / dt, 3), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"first {n} frames of the workload, FP32 twin + FP64 shadow, OpenMP {cores} threads "
                      f"(hardware_concurrency {os.cpu_count()}), {dt:.1f} s"}
    errs = {}
    rk, rs, rw, r64, rw64 = ref.export()
    storages = [("fp32", {})]
    if not no_f16:
        storages += [("fp16", dict(storage=2, weight_cap=255)), ("fp16_plain", dict(storage=1, weight_cap=255))]
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
        vol.close()
    return base, errs


if __name__ == "__main__":
    main()
