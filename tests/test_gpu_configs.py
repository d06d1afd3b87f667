"""Parity on the uneven, learned and combined workloads at their full size.

* mixed64: 64 views of 640x480 .. 2560x1440 (f = 0.86 W), K drawn from
  [4, 12] per view (scene.uneven_jobs), 8 sweeps, one batched schedule.
  Oracle row bands (exact: a pass reads only the rows next to a pixel, so
  a band with a 16-row margin is exact inside after 16 passes) of a
  2560x1440 view and of views at K = 10, 11 and 12 are compared bit for bit;
  the per-view launch schedule must give the batched result bit for bit on
  every view; every pixel of every view satisfies the gate's closed form,
  unit normals and the depth range.
* room1080 learned (20 x 1920x1080, K = 8): the learned initialisation +
  2 sweeps with a confidence threshold no pixel reaches is the plain
  2-sweep run bit for bit; at confidence 0.5 the full maps satisfy the
  same properties and agree with the 8-sweep run at least as well as two
  plain sweeps.
* orbit100 (100 x 1920x1080 on a 360-degree orbit, K = 10, depth range
  [1, 20]): oracle row bands of the plain 8-sweep run; the learned + 2-sweep
  run's properties; fusion of the learned depth maps against the oracle's
  fusion of the same maps, point for point.
"""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

_CACHE = {}


def _say(msg):
    print(msg, flush=True)


def _scene(name):
    if name not in _CACHE:
        from rustcolmap_amd.mvs import ViewData
        from rustcolmap_amd.scene import config_spec, render_scene
        _say("rendering %s" % name)
        sc = render_scene(config_spec(name), device="cuda", with_gt=False)
        _CACHE[name] = (sc, [ViewData.from_scene_view(v) for v in sc.views])
    return _CACHE[name]


def _oracle_band(views, job, opts, y0, rows=12):
    """The oracle's depth / cost / normals of rows [y0, y0 + rows) of job's
    reference view (computed on a band with a 2*sweeps-row margin)."""
    import oracle
    ref, srcs, zmin, zmax = job
    v = views[ref]
    ov = oracle.OracleViews([v] + [views[s] for s in srcs])
    P = 2 * opts.sweeps
    a, b = max(0, y0), min(v.h, y0 + rows)
    d, c, n, _, _ = ov.estimate_depth(0, list(range(1, 1 + len(srcs))), zmin, zmax, opts,
                                      rows=(max(0, a - P), min(v.h, b + P)))
    return a, b, d, c, n


def _check_band(dm, views, job, opts, y0):
    a, b, d, c, n = _oracle_band(views, job, opts, y0)
    ref = job[0]
    assert np.array_equal(bits(dm.depth[a:b]), bits(d[a:b])), (ref, a, b)
    assert np.array_equal(bits(dm.cost[a:b]), bits(c[a:b])), (ref, a, b)
    assert np.array_equal(bits(dm.normals[a:b]), bits(n[a:b])), (ref, a, b)


def _properties(ctx, dm, ref, zmin, zmax, cost_thr=0.38):
    tex = np.bitwise_count(ctx.view_census(ref)).astype(np.int32)
    z = dm.depth
    valid = z > 0
    assert np.all(valid <= ((tex >= 4) & (tex <= 44) & (dm.cost < np.float32(cost_thr)))), ref
    assert np.all((z[valid] > np.float32(zmin) * np.float32(0.5)) & (z[valid] < np.float32(zmax) * np.float32(2.0)))
    nn = np.linalg.norm(dm.normals.astype(np.float64), axis=-1)
    assert np.all(np.abs(nn - 1.0) < 1e-5), ref
    assert np.all((dm.cost >= 0) & (dm.cost <= 1.0)), ref
    return float(valid.mean())


def _agree(ref, other, tol=0.01):
    m = ref > 0
    o = other[m]
    ok = (o > 0) & (np.abs(1.0 / np.where(o > 0, o, 1.0) - 1.0 / ref[m]) <= tol / ref[m])
    return float(ok.mean()) if m.any() else float("nan")


def test_mixed64_uneven_batch(ctx):
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import uneven_jobs
    sc, views = _scene("mixed64")
    jobs = uneven_jobs(sc)
    ks = [len(j[1]) for j in jobs]
    assert min(ks) == 4 and max(ks) == 12
    opts = MvsOptions(num_sources=12, sweeps=8)
    ctx.upload_views(views)
    _say("mixed64 batched schedule (%d views, %.1f Mpx)" % (len(views), sum(v.w * v.h for v in views) / 1e6))
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    # views: 2560x1440 at K = 11 and 8, 1080p at K = 12 and 10, 720p and 480p at K = 12
    for ref, y0 in ((32, 0), (32, 700), (0, 1440 - 12), (27, 530), (17, 300), (58, 355), (42, 480 - 12)):
        assert (views[ref].w, len(jobs[ref][1])) in ((2560, 11), (2560, 8), (1920, 12), (1920, 10), (1280, 12),
                                                     (640, 12)), ref
        _say("mixed64 oracle band view %d (%dx%d, K=%d) rows %d+12" % (ref, views[ref].w, views[ref].h,
                                                                      len(jobs[ref][1]), y0))
        _check_band(out[ref], views, jobs[ref], opts, y0)
    valid = [_properties(ctx, out[r], r, 1.0, 10.0) for r in range(len(views))]
    _say("mixed64 valid fraction %.4f" % np.mean(valid))
    _say("mixed64 per-view launch schedule")
    per_view = ctx.estimate_depth_batch(jobs, opts, mode=1)
    for r in range(len(views)):
        for f in ("depth", "cost", "normals"):
            assert np.array_equal(bits(getattr(per_view[r], f)), bits(getattr(out[r], f))), (r, f)


def _net(ctx):
    from rustcolmap_amd.learned import DEFAULT_WEIGHTS, load_weights
    from rustcolmap_amd.mvs import ProposalNet
    net = ProposalNet(ctx, seed=0)
    load_weights(net, DEFAULT_WEIGHTS)
    return net


def test_room1080_learned_init(ctx):
    from rustcolmap_amd.mvs import MvsOptions, estimate_depth_learned
    from rustcolmap_amd.scene import select_sources
    sc, views = _scene("room1080")
    srcs = select_sources(sc.views, 8)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    ctx.upload_views(views)
    net = _net(ctx)
    short = MvsOptions(num_sources=8, sweeps=2)
    _say("room1080 plain 2 and 8 sweeps")
    plain2 = ctx.estimate_depth_batch(jobs, short, mode=0)
    plain8 = ctx.estimate_depth_batch(jobs, MvsOptions(num_sources=8, sweeps=8), mode=0)
    _say("room1080 learned, threshold above every confidence")
    none, acc = estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=2.0)
    assert acc == 0
    for r in range(len(views)):
        for f in ("depth", "cost", "normals"):
            assert np.array_equal(bits(getattr(none[r], f)), bits(getattr(plain2[r], f))), (r, f)
    _say("room1080 learned, confidence 0.5")
    learned, acc = estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=0.5)
    npix = sum(v.w * v.h for v in views)
    assert 0 < acc < npix
    for r in range(len(views)):
        _properties(ctx, learned[r], r, 1.0, 10.0)
    a_l = np.mean([_agree(plain8[r].depth, learned[r].depth) for r in range(len(views))])
    a_2 = np.mean([_agree(plain8[r].depth, plain2[r].depth) for r in range(len(views))])
    _say("room1080 agreement with 8 plain sweeps: learned %.4f, plain 2 sweeps %.4f, accepted %.4f"
         % (a_l, a_2, acc / npix))
    assert a_l > a_2
    net.close()


def test_orbit100_plain_learned_fusion(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions, estimate_depth_learned
    from rustcolmap_amd.scene import select_sources
    sc, views = _scene("orbit100")
    K, zmax = 10, 20.0
    srcs = select_sources(sc.views, K)
    jobs = [(r, srcs[r], 1.0, zmax) for r in range(len(views))]
    ctx.upload_views(views)
    plain = MvsOptions(num_sources=K, sweeps=8)
    _say("orbit100 plain 8 sweeps (100 views, K=10)")
    out = ctx.estimate_depth_batch(jobs, plain, mode=0)
    for ref, y0 in ((0, 0), (50, 540), (99, 1080 - 12)):
        _say("orbit100 oracle band view %d rows %d+12" % (ref, y0))
        _check_band(out[ref], views, jobs[ref], plain, y0)
    for r in range(len(views)):
        _properties(ctx, out[r], r, 1.0, zmax)
    del out
    net = _net(ctx)
    short = MvsOptions(num_sources=K, sweeps=2)
    _say("orbit100 learned + 2 sweeps, confidence 0.5")
    learned, acc = estimate_depth_learned(ctx, net, jobs, short, vol_k=4, conf_thr=0.5)
    assert acc > 0
    for r in range(len(views)):
        _properties(ctx, learned[r], r, 1.0, zmax)
    order = select_sources(sc.views, 2 * K)
    cell = 0.005
    _say("orbit100 fusion of the learned maps")
    got = ctx.fuse_resident([True] * len(views), order, cell, short)
    n_cand = ctx.stats().n_candidates
    ov = oracle.OracleViews(views, with_census=False)
    _say("orbit100 oracle fusion")
    exp, ncand_o, _ = ov.fuse([learned[r].depth for r in range(len(views))], order, short.min_consistent,
                              short.depth_diff, cell)
    assert n_cand == ncand_o
    assert len(got) == len(exp) and len(got) > 10000
    for f in ("x", "y", "z"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f
    assert np.array_equal(got["rgb"], exp["rgb"])
    net.close()
