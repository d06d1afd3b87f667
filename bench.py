"""Benchmark of the dense stereo + fusion path on one MI355X.

Workload (default): 20 views x 1920x1080, f = 1650, K = 8 support views,
8 sweeps (16 passes), 48 init hypotheses, inverse depth [0.1, 1.0], voxel
0.005 -- the corner room with 12 seeded objects, a textureless patch and
sigma = 2 DN noise (rustcolmap_amd.scene "room1080"), synthetic data.

A step = the whole stereo stage of the batch (init sweep + 16 passes + gate
for all 20 views) with the views' census/rays already resident in HBM.
value = depth Mpix.iter/s = sum(W*H) * sweeps / t(step); init and gate time
are inside t(step).  Timed steps launch directly on the library's stream
with HIP events around each stage (mode | 4), so the pass kernel's average
launch time and the exact evaluation counts come from the timed steps
themselves; the pass kernel's algorithmic bytes and roofline fraction follow
(bytes per pixel and pass = 120 + n_eval * (400 + 200 K)).  roofline.traffic
is the copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of the same kernel
on this workload, read from the committed summary under profiles/ (null when
the tuned kernel is not in it).  `--graph` times graph replays instead.
Fusion is timed for fused points/s and views/s.  The CPU baseline is the
oracle on a bounded row band of one view, on the host's cores.  The same
line carries `sfm_matching`: the second hot path (SfM descriptor matching,
20 frames x 2048 descriptors, all 190 pairs) measured as `--match` does,
with its own roofline and CPU baseline (`--no-match` skips it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# VALU issue peak: 256 CUs x 128 lanes per clock x 2.4 GHz = 78.6 T lane-ops/s.  The published MI355X FP32
# vector rate, 157.3 TFLOP/s, is this rate with an FMA counted as two operations; the pass kernel's own SQ
# counters agree (SQ_ACTIVE_INST_VALU ~ SQ_INSTS_VALU issue cycles at two clocks per wave64 instruction).
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
MFMA_BF16_PEAK_TFLOPS = 2500.0             # dense bf16 MFMA
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pass_kernel_room1080_traffic.txt")
KERNEL_SUMMARY = os.path.join(ROOT, "profiles", "stereo_room1080_kernel_trace.txt")
FUSION_SUMMARY = os.path.join(ROOT, "profiles", "fusion_room1080_counters.txt")
LEARNED_SUMMARY = os.path.join(ROOT, "profiles", "learned_room1080_kernel_trace.txt")
FP64_PEAK_TFLOPS = 78.6                    # FP64 vector
HBM_PEAK_GBS = 8000.0
# every kernel a fusion run launches (csrc/fuse.hip)
FUSION_KERNELS = ("mvs::valid_kernel", "mvs::cand_kernel", "mvs::key_init_kernel", "mvs::key_kernel", "mvs::insert_kernel",
                  "mvs::reduce_kernel", "mvs::compact_kernel", "mvs::exscan_kernel", "mvs::block_sum_kernel",
                  "mvs::scan_bsum_kernel", "mvs::slot_count_kernel", "mvs::slot_place_kernel", "mvs::fill_u32_k",
                  "mvs::fill_u64_k", "mvs::fill_u8_k")


def profiled_traffic(kernel, path=TRAFFIC_SUMMARY):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, each scaled
    by the copy-kernel calibration) from a committed summary written by
    --rocprof-summary; None when the file or the kernel is absent."""
    if not os.path.exists(path):
        return None
    fetch = write = None
    with open(path) as f:
        for line in f:
            if not line.startswith(kernel + " ") or "->" not in line:
                continue
            val = float(line.split("->")[1].split()[0])
            if " FETCH_SIZE " in line:
                fetch = val
            elif " WRITE_SIZE " in line:
                write = val
    if fetch is None or write is None:
        return None
    return fetch + write, os.path.relpath(path, ROOT)


def profiled_counter(kernel, counter, path):
    """Average per dispatch of `counter` for `kernel` from a committed
    --rocprof-summary file; None when the file or the line is absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(kernel + " ") and (" %s " % counter) in line and " avg " in line:
                return float(line.split(" avg ")[1].split()[0]), os.path.relpath(path, ROOT)
    return None


def profiled_stage_total(kernels, counter, path):
    """Sum over the kernels named in `kernels` (template arguments aside) of
    dispatches x per-dispatch value of `counter` (the copy-calibrated bytes
    where the summary gives them) in a committed --rocprof-summary file;
    None when the file or the counter is absent."""
    if not os.path.exists(path):
        return None
    total, seen = 0.0, 0
    tag = " %s " % counter
    with open(path) as f:
        for line in f:
            if tag not in line or "dispatches" not in line:
                continue
            if line.split(tag)[0].strip().split("<")[0] not in kernels:
                continue
            n = int(line.split("dispatches")[1].split()[0])
            val = float(line.split("->")[1].split()[0]) if "->" in line else float(line.split(" avg ")[1].split()[0])
            total += n * val
            seen += 1
    return (total, os.path.relpath(path, ROOT)) if seen else None


def profiled_kernel_avg_ns(kernel, path):
    """Average duration (ns) of `kernel` from the kernel table of a committed
    --rocprof-summary file (kernel, calls, total_ns, avg_ns, pct)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(kernel + " ") and "dispatches" not in line:
                parts = line[len(kernel):].split()
                if len(parts) >= 4:
                    return float(parts[2])
    return None


def fusion_fractions(ms_fuse, path=FUSION_SUMMARY):
    """HBM and FP64 rates of one fusion run: rocprof totals over every fusion
    kernel (FETCH_SIZE + WRITE_SIZE copy-calibrated bytes, and
    SQ_INSTS_VALU_FLOPS_FP64) of the committed profile of the same maps, over
    the measured fusion stage time."""
    fetch = profiled_stage_total(FUSION_KERNELS, "FETCH_SIZE", path)
    write = profiled_stage_total(FUSION_KERNELS, "WRITE_SIZE", path)
    fp64 = profiled_stage_total(FUSION_KERNELS, "SQ_INSTS_VALU_FLOPS_FP64", path)
    if ms_fuse <= 0 or fetch is None or write is None:
        return None
    b = fetch[0] + write[0]
    out = {"hbm": {"bytes_per_run": int(b), "GBs": round(b / (ms_fuse / 1000.0) / 1e9, 1), "peak_GBs": HBM_PEAK_GBS,
                   "frac": round(b / (ms_fuse / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)},
           "fp64": None, "source": fetch[1],
           "definition": "rocprof totals over every fusion kernel of one fusion of the same maps, over the "
                         "measured fusion stage time"}
    if fp64 is not None:
        out["fp64"] = {"flops_per_run": int(fp64[0]),
                       "TFLOPs": round(fp64[0] / (ms_fuse / 1000.0) / 1e12, 3), "peak_TFLOPs": FP64_PEAK_TFLOPS,
                       "frac": round(fp64[0] / (ms_fuse / 1000.0) / 1e12 / FP64_PEAK_TFLOPS, 5)}
    return out


def conv_kernel_mfma(path=LEARNED_SUMMARY):
    """MFMA rate of the network's convolution kernel from the committed
    learned-path profile: SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 flops per
    dispatch over its average dispatch time, against the dense bf16 peak."""
    kernel = "mvs::net::conv3x3_lds_kernel"
    mops = profiled_counter(kernel, "SQ_INSTS_VALU_MFMA_MOPS_BF16", path)
    ns = profiled_kernel_avg_ns(kernel, path)
    if mops is None or not ns:
        return None
    flops = mops[0] * 512.0
    t = flops / (ns * 1e-9) / 1e12
    return {"kernel": kernel, "mfma_flops_per_dispatch": int(flops), "avg_ns": ns, "achieved_TFLOPs": round(t, 1),
            "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS, "frac": round(t / MFMA_BF16_PEAK_TFLOPS, 4), "source": mops[1]}


def valu_roofline(kernel, launch_ms, path=KERNEL_SUMMARY):
    """The bound that actually limits the stereo kernels: wave-level VALU
    instructions per launch (SQ_INSTS_VALU from the committed profile of the
    same kernel and workload; the count is fixed by the seeded inputs) x 64
    lanes over the live launch time, against VALU_PEAK_TOPS."""
    c = profiled_counter(kernel, "SQ_INSTS_VALU", path)
    if c is None or launch_ms <= 0:
        return None
    achieved = c[0] * 64.0 / (launch_ms / 1000.0) / 1e12
    return {"instructions_per_launch": c[0], "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TOPS, 2),
            "unit": "T lane-ops/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "source": c[1]}


STALLS_SUMMARY = os.path.join(ROOT, "profiles", "pass_kernel_room1080_stalls.txt")


def lane_utilisation(kernel, path=STALLS_SUMMARY):
    """Active lanes per issued VALU instruction / 64 (SQ_THREAD_CYCLES_VALU
    over SQ_ACTIVE_INST_VALU x 64) from the committed profile; None when
    absent."""
    t = profiled_counter(kernel, "SQ_THREAD_CYCLES_VALU", path)
    a = profiled_counter(kernel, "SQ_ACTIVE_INST_VALU", path)
    if t is None or a is None or a[0] <= 0:
        return None
    return round(t[0] / (a[0] * 64.0), 4)


def stereo_roofline(pass_kernel, avg_pass_ms, traffic, floor, alg_gbs, alg_bytes_per_launch, time_source):
    """roofline object of the default line: the pass kernel against the VALU
    issue peak (the bound that limits it), with its measured HBM traffic next
    to the unique-bytes floor and the algorithmic byte rate of the
    measurement plan as separately labelled fields."""
    valu = valu_roofline(pass_kernel, avg_pass_ms)
    t = None if traffic is None else traffic[0]
    return {"bound": "valu", "kernel": pass_kernel,
            "achieved": None if valu is None else valu["achieved"],
            "peak": round(VALU_PEAK_TOPS, 2), "unit": "T lane-ops/s",
            "frac": None if valu is None else valu["frac"],
            "traffic": None if t is None else int(t),
            "avg_launch_ms": round(avg_pass_ms, 4),
            "valu_instructions_per_launch": None if valu is None else valu["instructions_per_launch"],
            "valu_source": None if valu is None else valu["source"],
            "lane_utilisation": lane_utilisation(pass_kernel),
            "peak_basis": "256 CUs x 128 lanes/clock x 2.4 GHz: the published 157.3 TFLOP/s FP32 vector rate with "
                          "an FMA counted as two operations",
            "hbm": {"measured_bytes": None if t is None else int(t),
                    "source": None if traffic is None else traffic[1],
                    "measured_GBs": None if t is None else round(t / (avg_pass_ms / 1000.0) / 1e9, 1),
                    "frac": None if t is None else round(t / (avg_pass_ms / 1000.0) / 1e9 / HBM_PEAK_GBS, 4),
                    "unique_bytes_floor": int(floor),
                    "traffic_over_floor": None if t is None else round(t / floor, 3),
                    "peak_GBs": HBM_PEAK_GBS},
            "algorithmic": {"GBs": round(alg_gbs, 1), "bytes_per_launch": int(alg_bytes_per_launch),
                            "definition": "120 B/px + n_eval x (400 + 200 K) B: every evaluated hypothesis with its "
                                          "full 5x5 window, as if every tap came from HBM (the early abort stops "
                                          "after ~18 of 200 taps, and windows come from LDS); a work rate, not an "
                                          "HBM rate"},
            "launch_time_source": time_source}


def host_threads():
    """Threads of the CPU baselines: every core this process may run on."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def pass_unique_bytes(views, jobs):
    """HBM floor of one pass launch: every byte it must touch, once -- the
    previous plane + cost read and the next written (2 x 20 B/px), the
    reference census and rays (16 B/px), and the census of every source
    view of every job (8 B/px), each read a single time."""
    total = 0
    for ref, srcs, _, _ in jobs:
        v = views[ref]
        total += v.w * v.h * 56
        total += sum(views[s].w * views[s].h * 8 for s in srcs)
    return total


# U-Net layer table of the proposal network (rustcolmap_amd/csrc/net.hip): (input channels, output channels,
# level); a level-l layer runs on the coarse grid subsampled by 2^l per axis
NET_LAYERS = [(64, 32, 0), (32, 32, 0), (32, 64, 1), (64, 64, 1), (64, 96, 2), (96, 96, 2), (96, 128, 3),
              (128, 128, 3), (128, 160, 4), (160, 160, 4), (288, 128, 3), (128, 128, 3), (224, 96, 2), (96, 96, 2),
              (160, 64, 1), (64, 64, 1), (96, 32, 0), (32, 32, 0)]


def net_conv_flops(views, stride=2):
    """Algorithmic MFMA work of one network forward over all views: 2 x 9 x
    cin x cout per output pixel of every 3x3 convolution, on each view's
    coarse grid (every stride-th pixel) at each layer's level."""
    total = 0
    for v in views:
        ch, cw = (v.h + stride - 1) // stride, (v.w + stride - 1) // stride
        for cin, cout, lev in NET_LAYERS:
            total += 2 * 9 * cin * cout * (ch >> lev) * (cw >> lev)
    return total


def rocprof_summary(path, copy_bytes=1 << 30, only=("mvs::", "mt::")):
    """Summarise rocprofv3 CSV output under `path`: per-kernel dispatch
    count / total / average duration (kernel_stats or kernel_trace CSV), and
    per-kernel averages of any collected counters (counter_collection CSV),
    with FETCH_SIZE / WRITE_SIZE scaled by what the same counter reads for
    the plain device copy of `copy_bytes` bytes that `--profile` runs."""
    import csv
    import glob
    from collections import defaultdict

    def short(name):
        return name.split("(")[0].replace("void ", "")[:90]

    lines = []
    for f in sorted(glob.glob(os.path.join(path, "**", "*kernel_stats.csv"), recursive=True)):
        lines.append("# kernel stats: %s" % os.path.relpath(f, path))
        lines.append("%-90s %8s %14s %12s %7s" % ("kernel", "calls", "total_ns", "avg_ns", "pct"))
        for r in csv.DictReader(open(f)):
            if only and not short(r["Name"]).startswith(only):
                continue
            lines.append("%-90s %8s %14s %12.0f %7.2f" % (short(r["Name"]), r["Calls"], r["TotalDurationNs"],
                                                          float(r["AverageNs"]), float(r["Percentage"])))
    for f in sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)):
        acc = defaultdict(lambda: [0.0, 0])
        for r in csv.DictReader(open(f)):
            key = (short(r["Kernel_Name"]), r["Counter_Name"])
            acc[key][0] += float(r["Counter_Value"])
            acc[key][1] += 1
        lines.append("# counters: %s" % os.path.relpath(f, path))
        copy = {}
        for (k, cn), (tot, cnt) in acc.items():
            if k.startswith("mvs::copy_kernel"):
                copy.setdefault(cn, []).append(tot / cnt)
        for (k, cn), (tot, cnt) in sorted(acc.items()):
            if only and not k.startswith(only):
                continue
            avg = tot / cnt
            line = "%-90s %-12s dispatches %5d  avg %.4g" % (k, cn, cnt, avg)
            if cn in ("FETCH_SIZE", "WRITE_SIZE") and copy.get(cn):
                # a plain copy of copy_bytes reads and writes copy_bytes each
                ref = max(copy[cn])
                scale = copy_bytes / ref if ref > 0 else float("nan")
                line += "  -> %.4g bytes (copy-calibrated x%.4g)" % (avg * scale, scale)
            lines.append(line)
    return "\n".join(lines)


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--rocprof-summary":
        print(rocprof_summary(sys.argv[2]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="room1080")
    ap.add_argument("--views", type=int, default=0, help="use only the first N views (0 = all)")
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sweeps", type=int, default=8)
    ap.add_argument("--cell", type=float, default=0.005)
    ap.add_argument("--mode", type=int, default=0, help="batch schedule (see mvs_estimate_depth_batch)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--graph", action="store_true", help="time HIP graph replays (stage times from an extra run)")
    ap.add_argument("--learned", action="store_true",
                    help="learned initialisation + 2 sweeps against the plain 8-sweep run (own JSON line)")
    ap.add_argument("--combined", action="store_true",
                    help="100 views 1080p on a 360-degree orbit, K=10: learned init + 2 sweeps + fusion (own JSON line)")
    ap.add_argument("--weights", default=None, help="proposal network weights (.npz)")
    ap.add_argument("--conf", type=float, default=0.5, help="confidence threshold of the learned initialisation")
    ap.add_argument("--conf-sweep", type=lambda v: [float(x) for x in v.split(",") if x], default=[0.3, 0.1],
                    help="further confidence thresholds reported beside --conf")
    ap.add_argument("--replace-sweep", type=lambda v: [float(x) for x in v.split(",") if x], default=[],
                    help="confidence thresholds at which the learned run is repeated with an accepted proposal "
                         "replacing the plain init sweep (instead of joining it as a 49th hypothesis)")
    ap.add_argument("--batched", action="store_true",
                    help="uneven batch: 64 views of mixed size, K uniform in [4, 12]; batched graph schedule vs "
                         "per-view launches (prints its own JSON line)")
    ap.add_argument("--match", action="store_true",
                    help="SfM descriptor matching: 20 frames x 2048 descriptors, all 190 pairs (prints its own "
                         "JSON line)")
    ap.add_argument("--match-frames", type=int, default=20)
    ap.add_argument("--match-features", type=int, default=2048)
    ap.add_argument("--match-exact", action="store_true", help="exact f32 tiles for every pair (no filter)")
    ap.add_argument("--no-match", action="store_true",
                    help="default run: skip the descriptor-matching block (sfm_matching)")
    ap.add_argument("--no-learned", action="store_true",
                    help="default run: skip the learned-initialisation block (learned)")
    ap.add_argument("--no-batched", action="store_true",
                    help="default run: skip the uneven-batch block (uneven_batch)")
    ap.add_argument("--profile", action="store_true",
                    help="one directly launched batch plus a 1 GiB device copy (for counter calibration), no JSON")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); multi-GPU runs are not part of it",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from rustcolmap_amd.mvs import MvsContext, MvsOptions, ViewData
    from rustcolmap_amd.scene import config_spec, render_scene, select_sources

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.match:
        return match_bench(args)
    if args.batched:
        return batched_bench(args)
    if args.learned or args.combined:
        return learned_bench(args)
    t0 = time.time()
    spec = config_spec(args.workload)
    if args.views:
        spec.n_views = args.views
    sc = render_scene(spec, device="cuda", with_gt=False)
    log("rendered %d views in %.1fs" % (len(sc.views), time.time() - t0))
    views = [ViewData.from_scene_view(v) for v in sc.views]
    K = min(args.k, len(views) - 1)
    srcs = select_sources(sc.views, K)
    opts = MvsOptions(num_sources=K, sweeps=args.sweeps)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    npix = sum(v.w * v.h for v in views)

    ctx = MvsContext(0)
    ctx.upload_views(views)
    # census time of a second upload: the first launch of a kernel in a process also waits for its code object to
    # load, which the census stage's events would otherwise include
    ctx.upload_views(views)
    ms_prepare = ctx.stats().ms_prepare

    if args.profile:
        ctx.estimate_depth_batch(jobs, opts, mode=(args.mode & 3) | 4, download=False)
        st = ctx.stats()
        log("profile run: passes %.1f ms, evals pass %d" % (st.ms_pass, st.n_eval_pass))
        pts = ctx.fuse_resident([True] * len(views), select_sources(sc.views, len(views) - 1), args.cell, opts)
        log("profile run: one fusion, %d points" % len(pts))
        from rustcolmap_amd import _lib as L
        L.check(L.load().mvs_debug_copy(ctx._h, 1 << 30), "mvs_debug_copy")   # mvs::copy_kernel, 1 GiB
        log("copy of 1073741824 bytes done")
        return

    tuned = None
    if not args.no_tune:
        t2 = time.time()
        tuned = ctx.autotune(jobs, opts, passes=1)
        log("autotune picked tile %dx%d persistent=%d evaluator=%d in %.1fs" % (tuned + (time.time() - t2,)))
        log("autotune candidates (est. ms per full stereo run): " +
            ", ".join("%s %.0f" % (n, t) for n, t in ctx.last_tune))

    mode = args.mode if args.graph else (args.mode & 3) | 4
    for _ in range(args.warmup):
        ctx.estimate_depth_batch(jobs, opts, mode=mode, download=False)
    torch.cuda.synchronize()
    acc = {"ms_init": 0.0, "ms_pass": 0.0, "ms_gate": 0.0, "n_eval_init": 0, "n_eval_pass": 0}
    step_walls = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        t2 = time.perf_counter()
        ctx.estimate_depth_batch(jobs, opts, mode=mode, download=False)
        step_walls.append(1000.0 * (time.perf_counter() - t2))
        if not args.graph:
            st = ctx.stats()
            for k_ in acc:
                acc[k_] += getattr(st, k_)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t1
    ms_step = 1000.0 * t_steps / args.steps
    value = npix * args.sweeps / (ms_step / 1000.0) / 1e6
    log("stereo step %.1f ms -> %.1f Mpix.iter/s" % (ms_step, value))
    if args.graph:
        # per-stage device times and exact evaluation counts (direct launches)
        ctx.estimate_depth_batch(jobs, opts, mode=(args.mode & 3) | 4, download=False)
        st = ctx.stats()
        nrun = 1
        for k_ in acc:
            acc[k_] = getattr(st, k_)
    else:
        nrun = args.steps
    ms_init, ms_pass, ms_gate = acc["ms_init"] / nrun, acc["ms_pass"] / nrun, acc["ms_gate"] / nrun
    n_eval_init, n_eval_pass = acc["n_eval_init"] // nrun, acc["n_eval_pass"] // nrun
    n_passes = 2 * args.sweeps
    b_eval = 400 + 200 * K
    pass_bytes = n_passes * npix * 120 + n_eval_pass * b_eval          # all passes of one step
    init_bytes = npix * 20 + n_eval_init * b_eval
    avg_pass_ms = ms_pass / n_passes
    achieved = pass_bytes / n_passes / (avg_pass_ms / 1000.0) / 1e9
    log("init %.1f ms, passes %.1f ms (%.2f ms/launch), gate %.2f ms; evals init %d pass %d"
        % (ms_init, ms_pass, avg_pass_ms, ms_gate, n_eval_init, n_eval_pass))
    st = ctx.stats()
    launch = (st.tile_w, st.tile_h, st.persistent, st.evaluator)
    pass_kernel = "mvs::%s<%d, %d, %d>" % (["pass_kernel", "pass_persistent_kernel", "pass_sorted_kernel"][launch[2]],
                                           launch[0], launch[1], launch[3])
    traffic = profiled_traffic(pass_kernel)

    # fusion of the resident depth maps
    order = select_sources(sc.views, len(views) - 1)
    pts = ctx.fuse_resident([True] * len(views), order, args.cell, opts)
    pts = ctx.fuse_resident([True] * len(views), order, args.cell, opts)
    sf = ctx.stats()
    fused_pps = len(pts) / (sf.ms_fuse / 1000.0) if sf.ms_fuse > 0 else 0.0
    views_per_s = len(views) / ((ms_prepare + ms_step + sf.ms_fuse) / 1000.0)
    log("fusion %.1f ms, %d candidates -> %d points" % (sf.ms_fuse, sf.n_candidates, len(pts)))

    learned = None
    if not args.no_learned:
        learned = learned_block(ctx, sc, views, jobs, K, opts, args, ms_prepare, ms_step, sf.ms_fuse)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(views, srcs, opts, args.cpu_seconds)

    init_kernel = ("mvs::init_sub_kernel<%d, %d>" % launch[:2] if launch[2] == 2 else
                   "mvs::init_kernel<%d, %d, %d>" % (launch[0], launch[1], launch[3]))
    init_valu = valu_roofline(init_kernel, ms_init)
    floor = pass_unique_bytes(views, jobs)

    out = {
        "metric": "depth Mpix.iter/s at 1080p K=8 (plain patch-match, whole stereo stage)",
        "value": round(value, 2),
        "unit": "Mpix*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u64",
        "data": "synthetic",
        "config": {"workload": "%s: %d views %dx%d f=%g K=%d sweeps=%d num_hyp=48 voxel=%g"
                   % (args.workload, len(views), views[0].w, views[0].h, views[0].cam.f, K, args.sweeps,
                      args.cell),
                   "global_batch": len(views), "parallelism": "single-gpu",
                   "launch": None if tuned is None else
                   {"tile": "%dx%d" % tuned[:2], "persistent": tuned[2], "evaluator": tuned[3],
                    "candidates_ms": {n: round(t, 1) for n, t in ctx.last_tune}}},
        "roofline": stereo_roofline(pass_kernel, avg_pass_ms, traffic, floor, achieved, pass_bytes / n_passes,
                                    "graph-free run after the timed steps" if args.graph
                                    else "HIP events around the pass launches of the timed steps"),
        "step_wall_ms": [round(x, 2) for x in step_walls],
        "stages_ms": {"census_rays": round(ms_prepare, 3), "init": round(ms_init, 3),
                      "passes": round(ms_pass, 3), "gate": round(ms_gate, 3),
                      "fusion": round(sf.ms_fuse, 3)},
        "mpix_iter_per_s_passes_only": round(npix * args.sweeps / (ms_pass / 1000.0) / 1e6, 2),
        "views_per_s": round(views_per_s, 3),
        "fused_points_per_s": round(fused_pps, 1),
        "fused_points": int(len(pts)),
        "fusion": fusion_fractions(sf.ms_fuse),
        "init": {"kernel": init_kernel, "ms": round(ms_init, 3),
                 "valu_frac": None if init_valu is None else init_valu["frac"],
                 "valu_achieved": None if init_valu is None else init_valu["achieved"],
                 "algorithmic_GBs": round(init_bytes / (ms_init / 1000.0) / 1e9, 1) if ms_init > 0 else None},
        "learned": learned,
        "cpu_baseline": cpu,
    }
    if not args.no_batched:
        # the uneven 64-view batch (mixed sizes up to 2560x1440, K in [4, 12]), batched graph schedule against
        # per-view launches, in the same run
        ctx.close()
        bargs = argparse.Namespace(**{**vars(args), "steps": 2, "warmup": 1, "views": 0})
        b = batched_bench(bargs, emit=False)
        out["uneven_batch"] = {k: b[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "schedules",
                                                 "speedup_batched_vs_per_view")}
    if not args.no_match:
        # the second hot path, SfM descriptor matching, measured in the same run
        margs = argparse.Namespace(**{**vars(args), "steps": 30, "warmup": 3, "match_exact": False,
                                      "profile": False})
        m = match_bench(margs, emit=False)
        out["sfm_matching"] = {k: m[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "matches",
                                                 "stage_ms", "roofline", "cpu_baseline") if k in m}
    print(json.dumps(out), flush=True)


def _agree(ref, other, tol=0.01):
    """Fraction of pixels valid in `ref` that are valid in `other` with
    inverse depth within `tol` (relative)."""
    m = ref > 0
    if not m.any():
        return float("nan")
    o = other[m]
    ok = (o > 0) & (np.abs(1.0 / np.where(o > 0, o, 1.0) - 1.0 / ref[m]) <= tol / ref[m])
    return float(ok.mean())


EQUAL_AGREEMENT_DEFINITION = ("plain views/s interpolated linearly in agreement between the plain runs of 1, 2, 3, "
                              "4 and 8 sweeps (plain_by_sweeps) at the learned run's agreement")


def _plain_vps_at(points, agree):
    """Views/s of the plain method at agreement `agree`, interpolated linearly
    between the (agreement, views/s) points of its runs at several sweep
    counts; the fastest run's rate when `agree` lies below all of them."""
    pts = sorted(points)
    for (a0, v0), (a1, v1) in zip(pts, pts[1:]):
        if a0 <= agree <= a1 and a1 > a0:
            return v0 + (v1 - v0) * (agree - a0) / (a1 - a0)
    if pts and agree < pts[0][0]:
        return max(v for _, v in pts)
    return None


def _gt_within(depth_maps, gt_maps, tol=0.01):
    """Against the renderer's depth: the share of valid pixels within `tol`
    (relative depth) of the truth, and the share of all foreground pixels
    that are valid and that close."""
    ok = val = fg = 0
    for r, d in depth_maps.items():
        g = gt_maps[r]
        m = (d > 0) & (g > 0)
        good = m & (np.abs(d - g) <= tol * np.where(g > 0, g, 1.0))
        ok += int(good.sum())
        val += int((d > 0).sum())
        fg += int((g > 0).sum())
    return {"of_valid": round(ok / max(val, 1), 4), "of_foreground": round(ok / max(fg, 1), 4)}


def learned_block(ctx, sc, views, jobs, K, opts, args, ms_prepare, ms_plain, ms_fuse_plain):
    """The second half of the headline metric on the same workload and
    context: learned initialisation + 2 sweeps (confidence 0.5, proposals as
    a 49th hypothesis) against the plain 8-sweep run just timed -- views/s
    of both (census + stereo + fusion), agreement of the learned maps with
    the plain 8-sweep maps (pixels valid there that are valid with inverse
    depth within 1 %), and the network stage's MFMA rate."""
    import torch
    from rustcolmap_amd.learned import DEFAULT_WEIGHTS, load_weights
    from rustcolmap_amd.mvs import MvsOptions, ProposalNet, estimate_depth_learned
    from rustcolmap_amd.scene import select_sources
    ref8 = {r: ctx.get_depth(r).depth for r in range(len(views))}   # the timed plain 8-sweep run's maps
    net = ProposalNet(ctx, seed=0)
    load_weights(net, DEFAULT_WEIGHTS)
    short = MvsOptions(num_sources=K, sweeps=2)
    conf = args.conf
    estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=conf, download=False)
    torch.cuda.synchronize()
    steps = max(1, args.steps)
    vol = netm = 0.0
    t = time.perf_counter()
    for _ in range(steps):
        _, acc = estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=conf, download=False)
        st = ctx.stats()
        vol += st.ms_learned_volume
        netm += st.ms_learned_net
    torch.cuda.synchronize()
    ms_learned = 1000.0 * (time.perf_counter() - t) / steps
    vol /= steps
    netm /= steps
    learned_d = {r: ctx.get_depth(r).depth for r in range(len(views))}
    order = select_sources(sc.views, len(views) - 1)
    # the first fusion of the learned maps also grows the candidate buffers (more valid pixels): time the second
    ctx.fuse_resident([True] * len(views), order, args.cell, short)
    ctx.fuse_resident([True] * len(views), order, args.cell, short)
    ms_fuse = ctx.stats().ms_fuse
    agree = float(np.mean([_agree(ref8[r], learned_d[r]) for r in ref8]))
    agree_both = _agree_valid_both(ref8, learned_d)
    vps_plain = len(views) / ((ms_prepare + ms_plain + ms_fuse_plain) / 1000.0)
    vps_learned = len(views) / ((ms_prepare + ms_learned + ms_fuse) / 1000.0)
    flops = net_conv_flops(views)
    net.close()
    # the plain method's own speed / accuracy curve on the same workload (1-4 sweeps, each timed with its
    # fusion), and its views/s interpolated at the learned run's agreement: the equal-accuracy comparison
    curve = {}
    mode = (args.mode & 3) | 4
    for ns in (1, 2, 3, 4):
        o = MvsOptions(num_sources=K, sweeps=ns)
        ctx.estimate_depth_batch(jobs, o, mode=mode, download=False)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            ctx.estimate_depth_batch(jobs, o, mode=mode, download=False)
        torch.cuda.synchronize()
        ms_s = 1000.0 * (time.perf_counter() - t) / steps
        d_s = {r: ctx.get_depth(r).depth for r in range(len(views))}
        ctx.fuse_resident([True] * len(views), order, args.cell, o)
        ms_f = ctx.stats().ms_fuse
        curve[str(ns)] = {"views_per_s": round(len(views) / ((ms_prepare + ms_s + ms_f) / 1000.0), 3),
                          "agreement_with_plain_8": round(float(np.mean([_agree(ref8[r], d_s[r]) for r in ref8])), 4),
                          "ms_stereo": round(ms_s, 2), "ms_fusion": round(ms_f, 2)}
    curve["8"] = {"views_per_s": round(vps_plain, 3), "agreement_with_plain_8": 1.0, "ms_stereo": round(ms_plain, 2),
                  "ms_fusion": round(ms_fuse_plain, 2)}
    vps_eq = _plain_vps_at([(c["agreement_with_plain_8"], c["views_per_s"]) for c in curve.values()], agree)
    out = {"views_per_s": round(vps_learned, 3), "plain_8_views_per_s": round(vps_plain, 3),
           "speedup": round(vps_learned / vps_plain, 3),
           "plain_views_per_s_at_equal_agreement": None if vps_eq is None else round(vps_eq, 3),
           "speedup_at_equal_agreement": None if vps_eq is None else round(vps_learned / vps_eq, 3),
           "equal_agreement_definition": EQUAL_AGREEMENT_DEFINITION,
           "plain_by_sweeps": curve,
           "agreement_with_plain_8": round(agree, 4),
           "agreement_definition": "pixels valid in the plain 8-sweep map that are valid with inverse depth "
                                   "within 1 % of it",
           "within_1pct_of_valid_in_both": round(agree_both, 4),
           "confidence_threshold": conf, "accepted_fraction": round(acc / sum(v.w * v.h for v in views), 4),
           "ms_stereo": round(ms_learned, 2), "ms_volume": round(vol, 2), "ms_net": round(netm, 2),
           "ms_fusion": round(ms_fuse, 2),
           "network": {"conv_flops": int(flops), "ms": round(netm, 2),
                       "achieved_TFLOPs": round(flops / (netm / 1000.0) / 1e12, 1) if netm > 0 else None,
                       "mfma_frac": round(flops / (netm / 1000.0) / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4)
                       if netm > 0 else None,
                       "definition": "2 x 9 x cin x cout per output pixel of the 18 3x3 convolutions, over the "
                                     "whole network stage time (convolutions, pooling, upsampling, head)",
                       "conv_kernel": conv_kernel_mfma()}}
    log("learned: %s" % out)
    return out


def _agree_valid_both(ref, other, tol=0.01):
    """Share of the pixels valid in both maps whose inverse depths agree
    within `tol` (relative)."""
    ok = tot = 0
    for r in ref:
        m = (ref[r] > 0) & (other[r] > 0)
        a, b = ref[r][m], other[r][m]
        ok += int((np.abs(1.0 / b - 1.0 / a) <= tol / a).sum())
        tot += int(m.sum())
    return ok / max(tot, 1)


def learned_bench(args):
    """Plain 8-sweep stereo (the reference run) vs the learned initialisation
    followed by 2 sweeps, on the same views; fusion after each."""
    import torch
    from rustcolmap_amd.learned import DEFAULT_WEIGHTS, load_weights
    from rustcolmap_amd.mvs import MvsContext, MvsOptions, ProposalNet, ViewData, estimate_depth_learned
    from rustcolmap_amd.scene import config_spec, render_scene, select_sources
    name = "orbit100" if args.combined else args.workload
    spec = config_spec(name)
    if args.views:
        spec.n_views = args.views
    K = 10 if args.combined else args.k
    zmax = 20.0 if args.combined else 10.0
    t0 = time.time()
    sc = render_scene(spec, device="cuda", with_gt=True)
    log("rendered %d views of %s in %.1fs" % (len(sc.views), name, time.time() - t0))
    views = [ViewData.from_scene_view(v) for v in sc.views]
    gt = {r: np.asarray(v.gt_depth, np.float32) for r, v in enumerate(sc.views)}
    K = min(K, len(views) - 1)
    srcs = select_sources(sc.views, K)
    jobs = [(r, srcs[r], 1.0, zmax) for r in range(len(views))]
    npix = sum(v.w * v.h for v in views)
    ctx = MvsContext(0)
    ctx.upload_views(views)
    ctx.upload_views(views)   # census time of a warm launch (see main)
    ms_prepare = ctx.stats().ms_prepare
    order = select_sources(sc.views, min(len(views) - 1, 2 * K))
    plain = MvsOptions(num_sources=K, sweeps=8)
    short = MvsOptions(num_sources=K, sweeps=2)
    if not args.no_tune:
        ctx.autotune(jobs, plain, passes=1)

    def timed(fn, steps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            r = fn()
        torch.cuda.synchronize()
        return r, 1000.0 * (time.perf_counter() - t) / steps

    def fuse():
        return ctx.fuse_resident([True] * len(views), order, args.cell, plain)

    steps = max(1, args.steps)
    # plain 8-sweep reference run
    _, ms_plain = timed(lambda: ctx.estimate_depth_batch(jobs, plain, mode=0, download=False), steps)
    ref8 = {r: ctx.get_depth(r).depth for r in range(len(views))}
    pts_plain, ms_fuse_plain = timed(fuse, 1)
    # plain method with 2 sweeps, no network
    _, ms_short = timed(lambda: ctx.estimate_depth_batch(jobs, short, mode=0, download=False), steps)
    short_d = {r: ctx.get_depth(r).depth for r in range(len(views))}
    # the plain method at other sweep counts: the speed / agreement curve the
    # learned initialisation has to beat
    by_sweeps = {}
    for ns in (1, 3, 4):
        o = MvsOptions(num_sources=K, sweeps=ns)
        _, ms_s = timed(lambda: ctx.estimate_depth_batch(jobs, o, mode=0, download=False), steps)
        d_s = {r: ctx.get_depth(r).depth for r in range(len(views))}
        by_sweeps[str(ns)] = {
            "views_per_s": round(len(views) / ((ms_prepare + ms_s + ms_fuse_plain) / 1000.0), 3),
            "agreement_with_plain_8": round(float(np.mean([_agree(ref8[r], d_s[r]) for r in ref8])), 4),
            "valid_fraction": round(float(np.mean([(d_s[r] > 0).mean() for r in d_s])), 4),
            "gt_within_1pct": _gt_within(d_s, gt), "ms_stereo": round(ms_s, 2)}
        log("plain %d sweeps: %s" % (ns, by_sweeps[str(ns)]))
    # learned initialisation + 2 sweeps, at each confidence threshold
    net = ProposalNet(ctx, seed=0)
    meta = load_weights(net, args.weights or DEFAULT_WEIGHTS)
    agree_s = float(np.mean([_agree(ref8[r], short_d[r]) for r in ref8]))
    valid = lambda d: float(np.mean([(d[r] > 0).mean() for r in d]))
    vps_plain = len(views) / ((ms_prepare + ms_plain + ms_fuse_plain) / 1000.0)
    vps_short = len(views) / ((ms_prepare + ms_short + ms_fuse_plain) / 1000.0)
    sweep = {}
    first = None
    runs = [(c, True) for c in [args.conf] + [c for c in args.conf_sweep if c != args.conf]]
    runs += [(c, False) for c in args.replace_sweep]
    replace_sweep = {}
    for conf, extra in runs:
        (_, acc), ms_learned = timed(lambda: estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=conf,
                                                                    download=False, extra=extra), steps)
        st = ctx.stats()
        learned_d = {r: ctx.get_depth(r).depth for r in range(len(views))}
        pts_learned, ms_fuse_learned = timed(fuse, 1)
        agree_l = float(np.mean([_agree(ref8[r], learned_d[r]) for r in ref8]))
        vps_learned = len(views) / ((ms_prepare + ms_learned + ms_fuse_learned) / 1000.0)
        row = {"views_per_s": round(vps_learned, 3), "speedup_vs_plain_8": round(vps_learned / vps_plain, 3),
               "agreement_with_plain_8": round(agree_l, 4), "valid_fraction": round(valid(learned_d), 4),
               "accepted_fraction": round(acc / npix, 4), "ms_stereo": round(ms_learned, 2),
               "ms_volume": round(st.ms_learned_volume, 2), "ms_net": round(st.ms_learned_net, 2),
               "ms_init_total": round(st.ms_init, 2), "ms_passes": round(st.ms_pass, 2),
               "ms_fusion": round(ms_fuse_learned, 2), "fused_points": int(len(pts_learned)),
               "gt_within_1pct": _gt_within(learned_d, gt)}
        log("conf >= %.2f%s: %s" % (conf, "" if extra else " (proposal replaces the sweep)", row))
        (sweep if extra else replace_sweep)["%.2f" % conf] = row
        if first is None:
            first = row
    curve = [(c["agreement_with_plain_8"], c["views_per_s"]) for c in by_sweeps.values()]
    curve += [(agree_s, vps_short), (1.0, vps_plain)]
    vps_eq = _plain_vps_at(curve, first["agreement_with_plain_8"])
    out = {"metric": "views/s, learned init + 2 sweeps vs plain 8 sweeps (%s)" % name,
           "value": first["views_per_s"], "unit": "views/s", "n_gpus": 1, "steps": steps, "warmup": 1,
           "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "%s: %d views %dx%d K=%d vol_k=4 conf>=%.2f zrange [1, %g]"
                      % (name, len(views), views[0].w, views[0].h, K, args.conf, zmax)},
           "plain_views_per_s": round(vps_plain, 3), "speedup_views_per_s": first["speedup_vs_plain_8"],
           "plain_views_per_s_at_equal_agreement": None if vps_eq is None else round(vps_eq, 3),
           "speedup_at_equal_agreement": None if vps_eq is None else round(first["views_per_s"] / vps_eq, 3),
           "equal_agreement_definition": EQUAL_AGREEMENT_DEFINITION,
           "agreement_with_plain_8": {"learned_2_sweeps": first["agreement_with_plain_8"],
                                      "plain_2_sweeps": round(agree_s, 4),
                                      "definition": "pixels valid in the 8-sweep plain map that are valid with "
                                                    "inverse depth within 1 %"},
           "plain_2_sweeps": {"views_per_s": round(vps_short, 3), "agreement_with_plain_8": round(agree_s, 4),
                              "valid_fraction": round(valid(short_d), 4), "ms_stereo": round(ms_short, 2),
                              "gt_within_1pct": _gt_within(short_d, gt)},
           "plain_8_sweeps": {"views_per_s": round(vps_plain, 3), "valid_fraction": round(valid(ref8), 4),
                              "ms_stereo": round(ms_plain, 2), "ms_census": round(ms_prepare, 2),
                              "ms_fusion": round(ms_fuse_plain, 2), "fused_points": int(len(pts_plain)),
                              "gt_within_1pct": _gt_within(ref8, gt)},
           "plain_by_sweeps": by_sweeps,
           "gt_definition": "renderer depth: share of valid pixels within 1 % relative depth (of_valid), share of "
                            "foreground pixels valid and that close (of_foreground)",
           "learned_by_confidence": sweep,
           "learned_replace_mode_by_confidence": replace_sweep,
           "weights": {k: (v.tolist() if hasattr(v, "tolist") else v) for k, v in meta.items()}}
    print(json.dumps(out), flush=True)


def batched_bench(args, emit=True):
    """64 views of uneven size (seeded draw over 640x480 .. 2560x1440,
    f = 0.86 W), K uniform in [4, 12] per view, 8 sweeps: one batched
    schedule (graph replay of one launch per stage) against one launch per
    view and stage.  With emit=False the result is returned, not printed."""
    import torch
    from rustcolmap_amd.mvs import MvsContext, MvsOptions, ViewData
    from rustcolmap_amd.scene import config_spec, render_scene, uneven_jobs
    spec = config_spec("mixed64")
    if args.views:
        spec.n_views = args.views
    t0 = time.time()
    sc = render_scene(spec, device="cuda", with_gt=False)
    log("rendered %d views in %.1fs" % (len(sc.views), time.time() - t0))
    views = [ViewData.from_scene_view(v) for v in sc.views]
    jobs = uneven_jobs(sc)
    ks = np.array([len(j[1]) for j in jobs])
    opts = MvsOptions(num_sources=12, sweeps=args.sweeps)
    npix = sum(v.w * v.h for v in views)
    ctx = MvsContext(0)
    ctx.upload_views(views)
    tuned = None if args.no_tune else ctx.autotune(jobs, opts, passes=1)
    res = {}
    for name, mode in (("batched_graph", 0), ("per_view_launches", 1)):
        for _ in range(args.warmup):
            ctx.estimate_depth_batch(jobs, opts, mode=mode, download=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ctx.estimate_depth_batch(jobs, opts, mode=mode, download=False)
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t1) / args.steps
        res[name] = {"ms_per_step": round(ms, 2), "mpix_iter_per_s": round(npix * args.sweeps / ms / 1000.0, 2)}
        log("%s: %.1f ms/step" % (name, ms))
    out = {"metric": "depth Mpix.iter/s, uneven batch (64 views, mixed sizes, K in [4,12])",
           "value": res["batched_graph"]["mpix_iter_per_s"], "unit": "Mpix*iter/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["batched_graph"]["ms_per_step"],
           "higher_is_better": True, "data": "synthetic",
           "config": {"workload": "mixed64: %d views, %.1f Mpx total, K %d..%d, sweeps %d"
                      % (len(views), npix / 1e6, int(ks.min()), int(ks.max()), args.sweeps),
                      "launch": None if tuned is None else
                      {"tile": "%dx%d" % tuned[:2], "persistent": tuned[2], "evaluator": tuned[3]}},
           "schedules": res,
           "speedup_batched_vs_per_view": round(res["per_view_launches"]["ms_per_step"]
                                                / res["batched_graph"]["ms_per_step"], 3)}
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


# f32 VALU operations/s without packed math: 256 CUs x 128 lanes/clock x 2.4 GHz (the published 157.3 TFLOP/s
# vector f32 rate counts an FMA as two)
MATCH_TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "match_filt_tile_match20x2048_traffic.txt")


def match_bench(args, emit=True):
    """Descriptor matching of run_sfm (sfm-lite/src/lib.rs:197-224):
    `--match-frames` frames of `--match-features` SIFT-shaped descriptors
    (synthetic sweep, 50 % overlap between neighbours), every pair i < j
    (exhaustive), ratio 0.75.  A step = one mvs_match_frames call over all
    pairs with the descriptors resident in HBM (match lists copied back to
    the host inside the step).  value = image pairs matched per second.
    With the filtered search (default) the dominant kernel is the MFMA tile
    kernel (one launch per step); its algorithmic work is 3 bf16 products x
    128 components x 2 flops per descriptor pair, against the dense bf16
    MFMA peak.  With --match-exact it is the
    exact distance-tile kernel: 3 f32 operations (difference, square, sum)
    per component, 384 per descriptor pair, against the VALU rate without
    FMA or packed math (the reference's rounding forbids both)."""
    import torch
    from rustcolmap_amd.matching import FrameMatcher, job_pairs, set_mode, synthetic_frames
    from rustcolmap_amd.mvs import MvsContext
    t0 = time.time()
    frames = synthetic_frames(args.match_frames, args.match_features, seed=0, overlap=0.5)
    pairs = job_pairs(len(frames), sequential=False)
    log("generated %d frames in %.1fs" % (len(frames), time.time() - t0))
    ctx = MvsContext(0)
    set_mode(ctx, args.match_exact)
    fm = FrameMatcher(ctx, frames)
    if args.profile:
        fm.match(pairs, 0.75)
        log("profile run: %r" % (fm.stats(),))
        from rustcolmap_amd import _lib as L
        L.check(L.load().mvs_debug_copy(ctx._h, 1 << 30), "mvs_debug_copy")   # mvs::copy_kernel, 1 GiB
        log("copy of 1073741824 bytes done")
        return
    for _ in range(max(args.warmup, 1)):
        fm.match(pairs, 0.75)
    torch.cuda.synchronize()
    stages = []
    call_ms = []
    step_ms = []
    import gc
    gc.collect()
    gc.disable()   # no collector pause of the interpreter inside the timed steps
    t1 = time.perf_counter()
    for _ in range(args.steps):
        t2 = time.perf_counter()
        res = fm.match(pairs, 0.75)
        step_ms.append(1000.0 * (time.perf_counter() - t2))
        call_ms.append(fm.last_call_ms)
        stages.append(fm.stats())
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t1) / args.steps
    gc.enable()
    st = stages[-1]
    n_match = sum(len(r) for r in res)
    distances = st["distances"]
    avg = {k: float(np.mean([x[k] for x in stages])) for k in ("exact_ms", "tile_ms", "collect_ms", "recheck_ms",
                                                               "select_ms")}
    host = {k: round(float(np.mean([x[k] for x in stages])), 4) for k in ("host_launch_ms", "host_wait_ms",
                                                                        "host_out_ms")}
    host["c_call_ms"] = round(float(np.mean(call_ms)), 4)
    host["python_step_ms"] = [round(x, 3) for x in step_ms]
    host["wait_out_ms_per_step"] = [(round(x["host_wait_ms"], 3), round(x["host_out_ms"], 3)) for x in stages]
    filtered = st["filtered_pairs"] == st["pairs"]
    if filtered:
        # dominant kernel: the MFMA tile kernel, one launch per step; 3 bf16
        # products of 128 components per distance
        kname = "mt::filt_tile_kernel<1>"
        launch_ms = avg["tile_ms"]
        flops = distances * 3 * 2 * 128.0
        peak, bound = MFMA_BF16_PEAK_TFLOPS, "mfma"
    else:
        kname = "mt::dist_tile_kernel"
        launch_ms = avg["exact_ms"]
        flops = distances * 384.0
        peak, bound = VALU_PEAK_TOPS, "valu"
    achieved = flops / (launch_ms / 1000.0) / 1e12
    if filtered:
        traffic = profiled_traffic(kname, MATCH_TRAFFIC_SUMMARY)
        if args.match_frames != 20 or args.match_features != 2048:
            traffic = None   # the committed summary is for the default workload
    else:
        traffic = None   # no committed counter summary for the exact tiles

    log("match: %.2f ms/step, stages %s, %d matches, %d overflow items" %
        (ms, ", ".join("%s %.3f" % kv for kv in avg.items()), n_match, st["overflow_items"]))
    out = {"metric": "SfM descriptor matching, image pairs/s (%d descriptors per frame, exhaustive pairs)"
                     % args.match_features,
           "value": round(len(pairs) / (ms / 1000.0), 2), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "match%dx%d: %d frames x %d SIFT-shaped descriptors, %d pairs, ratio 0.75"
                                  % (args.match_frames, args.match_features, len(frames), args.match_features,
                                     len(pairs)),
                      "search": "filtered" if filtered else "exact"},
           "matches": n_match, "overflow_items": st["overflow_items"],
           "stage_ms": {k: round(v, 4) for k, v in avg.items()}, "host_ms": host,
           "roofline": {"bound": bound, "achieved": round(achieved, 3), "peak": round(peak, 2),
                        "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                        "traffic": None if traffic is None else int(traffic[0]),
                        "flops_per_launch": int(flops), "avg_launch_ms": round(launch_ms, 4), "kernel": kname}}
    if not args.no_cpu:
        out["cpu_baseline"] = match_cpu_baseline(frames, pairs, args.cpu_seconds)
    ctx.close()
    if emit:
        print(json.dumps(out), flush=True)
    return out


def match_cpu_baseline(frames, pairs, seconds):
    """The oracle's match_pair (scalar C, both passes) over the same pairs
    in turn, one pair per host thread, until about `seconds` of wall time."""
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    ncores = host_threads()
    oracle.load()
    done = 0
    t = time.perf_counter()
    with ThreadPoolExecutor(ncores) as ex:
        while done < len(pairs) and time.perf_counter() - t < seconds:
            batch = pairs[done:done + ncores]
            list(ex.map(lambda p: oracle.match_pair(frames[p[0]], frames[p[1]], 0.75), batch))
            done += len(batch)
    el = time.perf_counter() - t
    val = done / el
    log("cpu baseline: %d pairs in %.1fs on %d threads -> %.3f pairs/s" % (done, el, ncores, val))
    return {"value": round(val, 4), "unit": "pairs/s", "cores": ncores, "kind": "port",
            "sample": "oracle match_pair on the first %d of the %d pairs, one pair per thread" % (done, len(pairs)),
            "seconds": round(el, 2)}


def cpu_baseline(views, srcs, opts, seconds):
    """Oracle (C restatement of the reference CPU path, rayon-like threads
    over rows) on whole reference views of the same batch in turn, then a
    row band of the next one, until about `seconds` of wall time; scaled to
    the same metric (pixels x sweeps / time)."""
    import oracle
    ncores = host_threads()

    def run(r, lo, hi):
        ov = oracle.OracleViews([views[r]] + [views[s] for s in srcs[r]])
        t = time.perf_counter()
        ov.estimate_depth(0, list(range(1, 1 + len(srcs[r]))), 1.0, 10.0, opts, nthreads=ncores, rows=(lo, hi))
        return time.perf_counter() - t

    v0 = views[0]
    probe = 2 * ncores
    per_row = run(0, v0.h // 2, v0.h // 2 + probe) / probe
    budget = seconds
    done_px = 0
    elapsed = 0.0
    parts = []
    r = 0
    while r < len(views) and elapsed < budget:
        v = views[r]
        rows = int(min(v.h, max(ncores, (budget - elapsed) / max(per_row, 1e-6))))
        rows -= rows % ncores if rows < v.h else 0
        lo = (v.h - rows) // 2
        dt = run(r, lo, lo + rows)
        elapsed += dt
        done_px += rows * v.w
        per_row = dt / rows
        parts.append("view %d rows %d-%d" % (r, lo, lo + rows) if rows < v.h else "view %d" % r)
        r += 1
    val = done_px * opts.sweeps / elapsed / 1e6
    log("cpu baseline: %s, %d threads, %.1fs -> %.4f Mpix.iter/s" % (", ".join(parts), ncores, elapsed, val))
    return {"value": round(val, 5), "unit": "Mpix*iter/s", "cores": ncores, "kind": "port",
            "sample": "oracle estimate_depth (init + %d passes, K=%d, %dx%d): %s"
                      % (2 * opts.sweeps, len(srcs[0]), v0.w, v0.h, ", ".join(parts)),
            "seconds": round(elapsed, 2)}


if __name__ == "__main__":
    main()
