"""Device arithmetic the kernels rely on for bit-exact parity with the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_fast_division_equals_ieee_division():
    """The integration kernels compute 1/z with the hardware reciprocal and one Newton step
    and the running mean's quotient with one Markstein correction; both must equal the IEEE
    division the oracle performs, bit for bit, over their whole input domains."""
    exe = os.path.join(HERE, "device_checks", "recip_exact")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run __graft_entry__.build()")
    print("device check: recip_exact (reciprocal over all normal floats, quotient over all a for b <= 256)", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip())
    assert r.returncode == 0 and r.stdout.count(" 0 mismatches") == 3, r.stdout + r.stderr
