"""bench.py without a GPU: the committed rocprofv3 summaries parse into the
roofline fields for the tilings they were taken at, and a multi-GPU request
is refused before anything touches a device."""

import glob
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

CELLS = 1024 ** 3


def profile_fields(path):
    fields = {}
    for line in open(path):
        if ":" in line and line.split(":")[0] in ("tile", "kernel"):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    return fields


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(ROOT, "profiles", "fused_step_kh1024_hbm_traffic*.md"))))
def test_traffic_profiles_parse_for_their_tiling(path):
    f = profile_fields(path)
    tile = tuple(int(v) for v in f["tile"].split(","))
    traffic, src = bench.measured_traffic(1024, 1024, 1024, f["kernel"], tile)
    assert src == os.path.relpath(path, ROOT)
    assert 40 * CELLS <= traffic <= 48 * CELLS  # algorithmic 40 B per cell plus halo re-reads
    valu = bench.measured_valu(1024, 1024, 1024, f["kernel"], tile)
    assert valu is not None and 0.0 < valu["busy"] <= 1.0 and 150 < valu["insts_per_cell"] < 400


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(ROOT, "profiles", "fused_step_kh1024_kernel_trace*.md"))))
def test_kernel_trace_profiles_parse_for_their_tiling(path):
    text = open(path).read()
    tile = tuple(int(v) for v in text.split("(tile_x,tile_y = ")[1].split(")")[0].split(","))
    kernel = "tf::k_line_wide<2, 16>"
    assert f"`{kernel}`" in text
    tr = bench.measured_trace(1024, 1024, 1024, kernel, tile)
    assert tr["source"] == os.path.relpath(path, ROOT)
    assert 5.0 < tr["avg_ms"] < 12.0  # 1024^3 steps: 42.95 GB at 3.6-8.6 TB/s


def test_profiles_of_another_tiling_or_grid_are_not_reported():
    assert bench.measured_traffic(1024, 1024, 1024, "tf::k_line_wide<2, 16>", (2, 7)) == (None, None)
    assert bench.measured_traffic(512, 512, 512, "tf::k_line_wide<2, 16>", (2, 128)) == (None, None)
    assert bench.measured_valu(1024, 1024, 1024, "tf::k_step_sims<8>", (2, 128)) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
