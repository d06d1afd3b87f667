"""Benchmark of the dense stereo + fusion path on one MI355X.

Workload (default): 20 views x 1920x1080, f = 1650, K = 8 support views,
8 sweeps (16 passes), 48 init hypotheses, inverse depth [0.1, 1.0], voxel
0.005 -- the corner room with 12 seeded objects, a textureless patch and
sigma = 2 DN noise (rustcolmap_amd.scene "room1080"), synthetic data.

A step = the whole stereo stage of the batch (init sweep + 16 passes + gate
for all 20 views), replayed from its captured HIP graph, with the views'
census/rays already resident in HBM.  value = depth Mpix.iter/s =
sum(W*H) * sweeps / t(step); init and gate time are inside t(step).

After the timed region: one directly launched run of the same batch gives
per-stage device times (HIP events on the library's stream) and the exact
evaluation counts, from which the pass kernel's algorithmic bytes and
roofline fraction follow (bytes per pixel and pass = 120 + n_eval * (400 +
200 K)); fusion is timed for fused points/s and views/s.  The CPU baseline
is the oracle on a bounded row band of one view, on the host's cores.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="room1080")
    ap.add_argument("--views", type=int, default=0, help="use only the first N views (0 = all)")
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--sweeps", type=int, default=8)
    ap.add_argument("--cell", type=float, default=0.005)
    ap.add_argument("--mode", type=int, default=0, help="batch schedule (see mvs_estimate_depth_batch)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
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
    ms_prepare = ctx.stats().ms_prepare

    if args.profile:
        ctx.estimate_depth_batch(jobs, opts, mode=(args.mode & 3) | 4, download=False)
        st = ctx.stats()
        log("profile run: passes %.1f ms, evals pass %d" % (st.ms_pass, st.n_eval_pass))
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda")   # 1 GiB
        y = torch.empty_like(x)
        x.fill_(1.0)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log("copy of 1073741824 bytes done")
        return

    tuned = None
    if not args.no_tune:
        t2 = time.time()
        tuned = ctx.autotune(jobs, opts, passes=1)
        log("autotune picked tile %dx%d persistent=%d evaluator=%d in %.1fs" % (tuned + (time.time() - t2,)))

    for _ in range(args.warmup):
        ctx.estimate_depth_batch(jobs, opts, mode=args.mode, download=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.estimate_depth_batch(jobs, opts, mode=args.mode, download=False)
    torch.cuda.synchronize()
    t_steps = time.perf_counter() - t1
    ms_step = 1000.0 * t_steps / args.steps
    value = npix * args.sweeps / (ms_step / 1000.0) / 1e6
    log("stereo step %.1f ms -> %.1f Mpix.iter/s" % (ms_step, value))

    # per-stage device times and exact evaluation counts (direct launches)
    ctx.estimate_depth_batch(jobs, opts, mode=(args.mode & 3) | 4, download=False)
    st = ctx.stats()
    n_passes = 2 * args.sweeps
    b_eval = 400 + 200 * K
    pass_bytes = n_passes * npix * 120 + st.n_eval_pass * b_eval          # all passes
    init_bytes = npix * 20 + st.n_eval_init * b_eval
    avg_pass_ms = st.ms_pass / n_passes
    achieved = pass_bytes / n_passes / (avg_pass_ms / 1000.0) / 1e9
    log("init %.1f ms, passes %.1f ms (%.2f ms/launch), gate %.2f ms; evals init %d pass %d"
        % (st.ms_init, st.ms_pass, avg_pass_ms, st.ms_gate, st.n_eval_init, st.n_eval_pass))

    # fusion of the resident depth maps
    order = select_sources(sc.views, len(views) - 1)
    pts = ctx.fuse_resident([True] * len(views), order, args.cell, opts)
    pts = ctx.fuse_resident([True] * len(views), order, args.cell, opts)
    sf = ctx.stats()
    fused_pps = len(pts) / (sf.ms_fuse / 1000.0) if sf.ms_fuse > 0 else 0.0
    views_per_s = len(views) / ((ms_prepare + ms_step + sf.ms_fuse) / 1000.0)
    log("fusion %.1f ms, %d candidates -> %d points" % (sf.ms_fuse, sf.n_candidates, len(pts)))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(views, srcs, opts, args.cpu_seconds)

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
                   {"tile": "%dx%d" % tuned[:2], "persistent": tuned[2], "evaluator": tuned[3]}},
        "roofline": {"bound": "hbm", "kernel": "pass_kernel", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_launch": int(pass_bytes / n_passes),
                     "avg_launch_ms": round(avg_pass_ms, 4)},
        "stages_ms": {"census_rays": round(ms_prepare, 3), "init": round(st.ms_init, 3),
                      "passes": round(st.ms_pass, 3), "gate": round(st.ms_gate, 3),
                      "fusion": round(sf.ms_fuse, 3)},
        "mpix_iter_per_s_passes_only": round(npix * args.sweeps / (st.ms_pass / 1000.0) / 1e6, 2),
        "views_per_s": round(views_per_s, 3),
        "fused_points_per_s": round(fused_pps, 1),
        "fused_points": int(len(pts)),
        "init_roofline_frac": round(init_bytes / (st.ms_init / 1000.0) / 1e9 / HBM_PEAK_GBS, 4)
        if st.ms_init > 0 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(views, srcs, opts, seconds):
    """Oracle (C restatement of the reference CPU path) on a row band of
    view 0, scaled to the same metric."""
    import oracle
    ncores = min(16, os.cpu_count() or 1)
    ov = oracle.OracleViews(views[: 1] + [views[s] for s in srcs[0]])
    local_srcs = list(range(1, 1 + len(srcs[0])))
    v = views[0]
    rows = 8
    t = time.perf_counter()
    ov.estimate_depth(0, local_srcs, 1.0, 10.0, opts, nthreads=ncores, rows=(v.h // 2, v.h // 2 + rows))
    dt = time.perf_counter() - t
    rows2 = int(min(v.h, max(rows, rows * seconds / max(dt, 1e-3))))
    rows2 = max(ncores, rows2 - rows2 % ncores)
    lo = max(0, v.h // 2 - rows2 // 2)
    hi = min(v.h, lo + rows2)
    t = time.perf_counter()
    ov.estimate_depth(0, local_srcs, 1.0, 10.0, opts, nthreads=ncores, rows=(lo, hi))
    dt = time.perf_counter() - t
    val = (hi - lo) * v.w * opts.sweeps / dt / 1e6
    log("cpu baseline: %d rows x %d px, %d threads, %.1fs -> %.4f Mpix.iter/s" % (hi - lo, v.w, ncores, dt, val))
    return {"value": round(val, 5), "unit": "Mpix*iter/s", "cores": ncores, "kind": "port",
            "sample": "oracle estimate_depth (init + %d passes, K=%d) on rows %d-%d of view 0 (%dx%d)"
                      % (2 * opts.sweeps, len(local_srcs), lo, hi, v.w, v.h),
            "seconds": round(dt, 2)}


if __name__ == "__main__":
    main()
