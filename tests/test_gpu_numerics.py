"""Device arithmetic the kernels rely on for bit-exact parity with the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reciprocal_equals_ieee_quotient_for_all_normal_depths():
    """The integration kernels compute 1/z with the hardware reciprocal and one Newton step;
    for every normal z < 2^126 it must equal 1.0f / z bit for bit (the oracle divides)."""
    exe = os.path.join(HERE, "device_checks", "recip_exact")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run __graft_entry__.build()")
    print("device check: recip_exact over all positive normal floats below 2^126", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0 and " 0 mismatches" in r.stdout, r.stdout + r.stderr
