"""Writes tests/golden/synth_room_small.json: SHA-256 digests of selected frames and poses of
the small-room sequence (seed 20240611), pinning the generator across hosts. Run from the
repository root after building libvh."""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

vh, wl = __graft_entry__.package()
scene = vh.Scene(1, wl.SEED)
out = {"preset": 1, "seed": wl.SEED, "frames": {}}
for f in (0, 1, 17, 39):
    d = scene.depth(f)
    T = scene.pose(f)
    out["frames"][str(f)] = {"depth_sha256": hashlib.sha256(d.tobytes()).hexdigest(),
                             "pose_sha256": hashlib.sha256(T.tobytes()).hexdigest(),
                             "valid_pixels": int((d > 0).sum()), "depth_sum_mm": int(d.astype(int).sum())}
json.dump(out, open(os.path.join(ROOT, "tests", "golden", "synth_room_small.json"), "w"), indent=1)
print("written")
