"""The committed rocprof summaries feed bench.py's roofline fields: the pass
and matching kernels' HBM traffic (copy-calibrated FETCH_SIZE + WRITE_SIZE)
and the pass kernel's VALU instruction count.  These checks keep the files in
the form bench.py parses, and the numbers consistent with each other."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

PASS = "mvs::pass_persistent_kernel<8, 8, 3>"
TILE = "mt::filt_tile_kernel<1>"


def _avg_ns(path, kernel):
    with open(path) as f:
        for line in f:
            if line.startswith(kernel + " ") and " avg " not in line:
                return float(line.split()[-2])
    return None


def test_pass_kernel_traffic_and_valu_parse():
    t = bench.profiled_traffic(PASS)
    assert t is not None and t[1] == os.path.join("profiles", "pass_kernel_room1080_traffic.txt")
    # well above the ~5 GB a launch must move at least (every source census once,
    # the reference maps), well below the 5.85e11 algorithmic bytes
    assert 5e9 < t[0] < 5e10
    v = bench.valu_roofline(PASS, 11.8)
    assert v is not None and v["unit"] == "T lane-ops/s"
    assert 1e9 < v["instructions_per_launch"] < 1e11
    assert 0.0 < v["frac"] < 1.0
    assert v["achieved"] == pytest.approx(v["instructions_per_launch"] * 64 / 11.8e-3 / 1e12, rel=1e-3)


def test_profiles_agree_on_the_pass_launch_time():
    """The kernel trace's average pass launch and the traffic file's stated
    launch time describe the same build (within a few percent)."""
    kt = os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt")
    ns = _avg_ns(kt, PASS)
    assert ns is not None and 5e6 < ns < 3e7
    with open(os.path.join(ROOT, "profiles", "pass_kernel_room1080_traffic.txt")) as f:
        text = f.read()
    ms = float(text.split("HBM bytes per launch in ")[1].split()[0])
    assert ms == pytest.approx(ns / 1e6, rel=0.05)


def test_matching_tile_kernel_traffic_parses():
    t = bench.profiled_traffic(TILE, bench.MATCH_TRAFFIC_SUMMARY)
    assert t is not None and 1e8 < t[0] < 1e10
    kt = os.path.join(ROOT, "profiles", "match_filt_tile_match20x2048_kernel_trace.txt")
    ns = _avg_ns(kt, TILE)
    assert ns is not None and 5e5 < ns < 5e6
