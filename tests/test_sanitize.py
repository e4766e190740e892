"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (g++, no GPU): the
synthetic frame generator, the settings / PNG / pose readers (also on malformed files) and
the CPU oracle, driven by tests/host_checks/sanitize_host.cpp."""
import os
import subprocess

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "voxel-hashing-sdf_amd")


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fopenmp", "-I", os.path.join(ROOT, "include"), "-I",
           os.path.join(PKG, "csrc"), "-I", os.path.join(ROOT, "oracle"), "-o", exe,
           os.path.join(ROOT, "tests", "host_checks", "sanitize_host.cpp"), os.path.join(PKG, "csrc", "vh_io.cpp"),
           os.path.join(PKG, "csrc", "vh_synth.cpp"), os.path.join(ROOT, "oracle", "vh_oracle.cpp"), "-lz",
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    yaml = os.path.join(ROOT, "tests", "golden", "settings_scannet.yaml")
    r = subprocess.run([exe, str(tmp_path), yaml], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-6000:]
