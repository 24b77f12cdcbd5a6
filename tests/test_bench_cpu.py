"""bench.py without a GPU: the committed rocprofv3 summaries parse into the
roofline fields for the tilings they were taken at, only while the kernel
sources are the ones they were taken on, and a multi-GPU request is refused
before anything touches a device."""

import glob
import json
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

CELLS = 1024 ** 3


def committed(stem):
    return sorted(glob.glob(os.path.join(ROOT, "profiles", stem + "*.md")))


def own_hash(monkeypatch, path):
    """Parse a committed summary as if the kernel sources were the ones it
    was taken on."""
    f = bench.profile_fields(open(path).read())
    assert len(f.get("source_hash", "")) == 16, f"{path} records no kernel source hash"
    monkeypatch.setattr(bench, "kernel_source_hash", lambda root=ROOT: f["source_hash"])
    return f["kernel"], tuple(int(v) for v in f["tile"].split(","))


@pytest.mark.parametrize("path", committed("fused_step_kh1024_hbm_traffic"))
def test_traffic_profiles_parse_for_their_tiling(path, monkeypatch):
    kernel, tile = own_hash(monkeypatch, path)
    traffic, src = bench.measured_traffic(1024, 1024, 1024, kernel, tile)
    assert src == os.path.relpath(path, ROOT)
    assert 40 * CELLS <= traffic <= 48 * CELLS  # algorithmic 40 B per cell plus halo re-reads


@pytest.mark.parametrize("path", committed("fused_step_kh1024_valu"))
def test_valu_profiles_parse_for_their_tiling(path, monkeypatch):
    kernel, tile = own_hash(monkeypatch, path)
    valu = bench.measured_valu(1024, 1024, 1024, kernel, tile)
    assert valu is not None and 0.0 < valu["busy"] <= 1.0 and 150 < valu["insts_per_cell"] < 400


@pytest.mark.parametrize("path", committed("fused_step_kh1024_kernel_trace"))
def test_kernel_trace_profiles_parse_for_their_tiling(path, monkeypatch):
    kernel, tile = own_hash(monkeypatch, path)
    assert f"`{kernel}`" in open(path).read()
    tr = bench.measured_trace(1024, 1024, 1024, kernel, tile)
    assert tr["source"] == os.path.relpath(path, ROOT)
    assert 5.0 < tr["avg_ms"] < 12.0  # 1024^3 steps: 42.95 GB at 3.6-8.6 TB/s


def test_profiles_of_another_tiling_grid_or_build_are_not_reported(tmp_path):
    """A summary is reported only for its grid, kernel and tiling, and only
    while the kernel sources hash to what it records."""
    for rel in bench.KERNEL_SOURCES:
        os.makedirs(os.path.dirname(tmp_path / rel), exist_ok=True)
        shutil.copy(os.path.join(ROOT, rel), tmp_path / rel)
    os.makedirs(tmp_path / "profiles")
    kernel = "tf::k_line_wide<2, 16>"

    def write(h):
        (tmp_path / "profiles" / "fused_step_kh1024_hbm_traffic.md").write_text(
            f"# t\n\ntraffic_bytes_per_launch: {42 * CELLS}\ntile: 2,128\nkernel: {kernel}\nsource_hash: {h}\n")

    write(bench.kernel_source_hash(str(tmp_path)))
    assert bench.measured_traffic(1024, 1024, 1024, kernel, (2, 128), root=str(tmp_path))[0] == 42 * CELLS
    assert bench.stale_profiles(str(tmp_path)) == []
    assert bench.measured_traffic(1024, 1024, 1024, kernel, (2, 7), root=str(tmp_path)) == (None, None)
    assert bench.measured_traffic(512, 512, 512, kernel, (2, 128), root=str(tmp_path)) == (None, None)
    assert bench.measured_traffic(1024, 1024, 1024, "tf::k_step_sims<8>", (2, 128), root=str(tmp_path)) == (None,
                                                                                                               None)
    # the kernel changes: the summary describes another build
    with open(tmp_path / bench.KERNEL_SOURCES[0], "a") as f:
        f.write("\n// changed\n")
    assert bench.measured_traffic(1024, 1024, 1024, kernel, (2, 128), root=str(tmp_path)) == (None, None)
    assert bench.stale_profiles(str(tmp_path)) == ["profiles/fused_step_kh1024_hbm_traffic.md"]


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
