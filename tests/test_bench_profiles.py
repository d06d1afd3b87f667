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

PASS = "mvs::pass_sorted_kernel<24, 8, 3>"
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


def _room1080_like():
    """views and jobs of the room1080 shape (20 x 1920x1080, K = 8) without
    rendering: the floor depends on sizes and source lists only."""
    from types import SimpleNamespace
    views = [SimpleNamespace(w=1920, h=1080) for _ in range(20)]
    jobs = [(r, [s for s in range(20) if s != r][:8], 1.0, 10.0) for r in range(20)]
    return views, jobs


def test_roofline_object_is_a_bound():
    """The default line's roofline: the pass kernel against the VALU issue
    peak, 0 < frac <= 1 at the committed profile's launch time; the measured
    HBM traffic is a fraction of the HBM peak and at least the unique-bytes
    floor; the algorithmic byte rate is a separate, labelled field."""
    ns = _avg_ns(os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt"), PASS)
    views, jobs = _room1080_like()
    floor = bench.pass_unique_bytes(views, jobs)
    assert floor == 20 * 1920 * 1080 * (56 + 8 * 8)
    r = bench.stereo_roofline(PASS, ns / 1e6, bench.profiled_traffic(PASS), floor, 49917.9, 5.85e11, "test")
    assert r["bound"] == "valu" and r["unit"] == "T lane-ops/s"
    assert 0.0 < r["frac"] <= 1.0
    assert r["achieved"] == pytest.approx(r["frac"] * r["peak"], rel=1e-3)
    assert 0.0 < r["hbm"]["frac"] <= 1.0
    assert r["hbm"]["traffic_over_floor"] >= 1.0
    assert r["algorithmic"]["GBs"] == 49917.9
    assert 0.5 < r["lane_utilisation"] <= 1.0
    # the init kernel on the same basis
    init = "mvs::init_sub_kernel<24, 8>"
    ns_i = _avg_ns(os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt"), init)
    v = bench.valu_roofline(init, ns_i / 1e6)
    assert v is not None and 0.0 < v["frac"] <= 1.0


def test_network_flops_from_the_layer_table():
    """net_conv_flops counts 2 x 9 x cin x cout per output pixel: at one
    level-0 pixel of a 2x2-subsampled 2x2 view every level > 0 layer sees
    0 pixels."""
    from types import SimpleNamespace
    tiny = [SimpleNamespace(w=2, h=2)]
    lvl0 = sum(2 * 9 * ci * co for ci, co, lev in bench.NET_LAYERS if lev == 0)
    assert bench.net_conv_flops(tiny) == lvl0
    assert bench.net_conv_flops([SimpleNamespace(w=1920, h=1080)]) > 1e11


def test_fusion_fractions_parse():
    """Fusion's HBM and FP64 fractions: rocprof totals over every fusion
    kernel of one run (the stereo fill kernel, mvs::fill_u64_kernel, is not
    one of them), over a stage time; both are fractions of their peaks."""
    fe = bench.profiled_stage_total(bench.FUSION_KERNELS, "FETCH_SIZE", bench.FUSION_SUMMARY)
    wr = bench.profiled_stage_total(bench.FUSION_KERNELS, "WRITE_SIZE", bench.FUSION_SUMMARY)
    assert fe is not None and wr is not None
    # more than the candidate kernel's own reads, less than a pass launch
    assert 7e8 < fe[0] + wr[0] < 5e9
    f = bench.fusion_fractions(2.8)
    assert f["hbm"]["bytes_per_run"] == int(fe[0] + wr[0])
    assert 0.0 < f["hbm"]["frac"] < 1.0 and 0.0 < f["fp64"]["frac"] < 1.0
    assert f["hbm"]["GBs"] == pytest.approx((fe[0] + wr[0]) / 2.8e-3 / 1e9, rel=1e-3)


def test_conv_kernel_mfma_from_counters():
    """The network's convolution kernel: MFMA flops per dispatch from
    SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512, which equals its MFMA instruction
    count x 16384 (v_mfma_f32_16x16x32_bf16) in the same profile."""
    c = bench.conv_kernel_mfma()
    assert c is not None and 0.0 < c["frac"] < 1.0
    n = bench.profiled_counter("mvs::net::conv3x3_lds_kernel", "SQ_INSTS_VALU_MFMA_BF16", bench.LEARNED_SUMMARY)
    assert n is not None
    assert c["mfma_flops_per_dispatch"] == pytest.approx(n[0] * 16384, rel=2e-3)


@pytest.mark.parametrize("shape", ["24, 8", "16, 8", "16, 16"])
def test_every_sorted_pass_shape_is_profiled(shape):
    """The tuner picks among the sorted pass shapes by timing; whichever it
    picks, the default line's roofline fields come from the committed
    profiles: HBM traffic, lane utilisation and VALU instructions, and a
    kernel-trace launch time within a few percent of the other shapes'."""
    k = "mvs::pass_sorted_kernel<%s, 3>" % shape
    t = bench.profiled_traffic(k)
    assert t is not None and 4.98e9 <= t[0] < 5e10
    lu = bench.lane_utilisation(k)
    assert lu is not None and 0.5 < lu <= 1.0
    v = bench.valu_roofline(k, 10.8)
    assert v is not None and 0.0 < v["frac"] < 1.0
    ns = _avg_ns(os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt"), k)
    ref = _avg_ns(os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt"), PASS)
    assert ns is not None and ns == pytest.approx(ref, rel=0.05)


def test_plain_rate_at_equal_agreement_interpolates_between_sweep_counts():
    # (agreement, views/s) of plain runs at 1, 2, 4 and 8 sweeps
    curve = [(0.80, 500.0), (0.82, 300.0), (0.86, 200.0), (1.0, 100.0)]
    assert bench._plain_vps_at(curve, 0.84) == pytest.approx(250.0)
    assert bench._plain_vps_at(curve, 0.86) == pytest.approx(200.0)
    # order of the points does not matter
    assert bench._plain_vps_at(curve[::-1], 0.93) == pytest.approx(150.0)
    # below every plain run: the fastest plain run already agrees better
    assert bench._plain_vps_at(curve, 0.5) == 500.0
    assert bench._plain_vps_at([], 0.9) is None
