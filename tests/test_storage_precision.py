"""Why the 4-byte format that meets a TSDF error of 1e-3 is F16R, not plain binary16: a numpy
model of one frame's storage round trip, on the capped running mean the integration keeps
(s <- s + (d - s) / (w + 1), weight cap 255).

Near |s| = 1 a binary16 value has an ulp of 2^-11 (4.9e-4). Once the weight reaches the cap,
an update moves s by (d - s) / 256. That is below half an ulp whenever |d - s| < 1/16, so
round-to-nearest stagnates and drifts with the noise. Stochastic rounding, seeded per voxel
and frame, removes the bias but turns the lost increments into a random walk of about
ulp * sqrt(w / 8). Neither keeps every voxel within 1e-3 of the FP64 mean. The F16R residual
(int8, in units of ulp/256) adds 8 significant bits and does.
"""
import numpy as np

N, FRAMES, CAP = 40000, 400, 255


def _round_sr(x, rng):
    """Stochastic rounding of FP32 values to binary16 (probability of rounding up = the
    fraction of the ulp)."""
    lo = x.astype(np.float16)
    lo = np.where(lo.astype(np.float32) > x, np.nextafter(lo, np.float16(-np.inf)), lo)
    hi = np.nextafter(lo, np.float16(np.inf))
    span = hi.astype(np.float32) - lo.astype(np.float32)
    p = np.where(span > 0, (x - lo.astype(np.float32)) / np.where(span > 0, span, 1), 0)
    return np.where(rng.random(len(x)) < p, hi, lo).astype(np.float32)


def _round_f16r(x):
    """F16R: binary16 value h plus an int8 residual r in units of ulp(h)/256."""
    h = x.astype(np.float16).astype(np.float32)
    ulp = np.spacing(np.abs(h).astype(np.float16)).astype(np.float32)
    r = np.clip(np.rint((x - h) / (ulp / 256)), -128, 127)
    return (h + r * (ulp / 256)).astype(np.float32)


def _run():
    rng = np.random.default_rng(3)
    true = rng.uniform(-1, 1, N)
    s64 = np.zeros(N)
    w = np.zeros(N)
    s = {k: np.zeros(N, np.float32) for k in ("rn", "sr", "f16r")}
    for _ in range(FRAMES):
        d = np.minimum(1.0, true + rng.normal(0, 0.2, N)).astype(np.float32)
        wn = w + 1
        inv = (1.0 / wn).astype(np.float32)
        s64 = s64 + (d.astype(np.float64) - s64) / wn
        upd = {k: v + (d - v) * inv for k, v in s.items()}
        s["rn"] = upd["rn"].astype(np.float16).astype(np.float32)
        s["sr"] = _round_sr(upd["sr"], rng)
        s["f16r"] = _round_f16r(upd["f16r"])
        w = np.minimum(wn, CAP)
    return {k: np.abs(v.astype(np.float64) - s64) for k, v in s.items()}, {k: v - s64 for k, v in s.items()}


def test_binary16_cannot_meet_1e3_and_f16r_does():
    err, signed = _run()
    rms = {k: float(np.sqrt((e ** 2).mean())) for k, e in err.items()}
    # plain binary16: max error an order above 1e-3 with either rounding
    assert err["rn"].max() > 5e-3 and err["sr"].max() > 5e-3
    # stochastic rounding: no bias, smaller RMS than round-to-nearest, still above 1e-3
    assert abs(signed["sr"].mean()) < 0.2 * abs(signed["rn"].mean())
    assert rms["sr"] < rms["rn"] and rms["sr"] > 1e-3
    # F16R: every voxel within 1e-4 of the FP64 mean
    assert err["f16r"].max() < 1e-4
