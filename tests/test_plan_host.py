"""The batch-plan builder (tinyfluids_amd/csrc/tf_plan.cpp) under address and
undefined-behaviour sanitizers, on the CPU: tests/host/plan_check.cpp is
built with the host compiler (no HIP, no device) and builds plans over
seeded uneven batches, the benchmark batch, 2^30-cell grids and bad
requests, checking that every output cell is written exactly once by the
kernels' own tile decoders and that every table stays inside the
workspace."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "tinyfluids_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_plan_builder_under_sanitizers(tmp_path):
    exe = str(tmp_path / "plan_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-Wall", "-Werror",
           "-I", CSRC, os.path.join(ROOT, "tests", "host", "plan_check.cpp"), os.path.join(CSRC, "tf_plan.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "plan_check: OK" in r.stdout
