"""Named workloads: scene preset + volume parameters.

room_small_1cm    40 frames, 4x3x2.5 m room + 3 spheres + 2 boxes, 1 cm voxels, 4 cm truncation,
                  2^18 buckets, FP32 (the CPU-runnable case).
room_medium_5mm   500 frames, 8x6x3 m room + 12 primitives, 5 mm voxels, 2 cm truncation,
                  2^21 buckets, 1.5M-block pool, FP32; marching cubes every 100 frames + final.
room_medium_5mm_f16  the same with binary16 TSDF + uint16 weight capped at 255.
floor_large_4mm   2000 frames, 20x20x4 m floor plan (2x2 rooms, 10 cm walls, doorways,
                  40 primitives), 4 mm voxels, 1.6 cm truncation, 2^23 buckets, 6M-block pool, FP16.
All share the 640x480 camera (fx = fy = 525, cx = 319.5, cy = 239.5), uint16 mm depth,
8^3 blocks and iso-surface 0. Depth beyond max_depth (5 m) is ignored.
"""

CAMERA = dict(width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, min_depth=0.1, max_depth=5.0)

SEED = 20240611  # scene layout seed; depth noise uses SEED + 1, pose errors SEED + 2

WORKLOADS = {
    "room_small_1cm": dict(CAMERA, scene=1, frames=40, vox_size=0.01, trunc=0.04, log2_buckets=18,
                           pool_blocks=1 << 18, storage=0, weight_cap=0, mc_every=0),
    "room_medium_5mm": dict(CAMERA, scene=2, frames=500, vox_size=0.005, trunc=0.02, log2_buckets=21,
                            pool_blocks=1_500_000, storage=0, weight_cap=0, mc_every=100),
    "room_medium_5mm_f16": dict(CAMERA, scene=2, frames=500, vox_size=0.005, trunc=0.02, log2_buckets=21,
                                pool_blocks=1_500_000, storage=1, weight_cap=255, mc_every=100),
    "floor_large_4mm": dict(CAMERA, scene=5, frames=2000, vox_size=0.004, trunc=0.016, log2_buckets=23,
                            pool_blocks=6_000_000, storage=1, weight_cap=255, mc_every=0, gn_iters=5),
}
