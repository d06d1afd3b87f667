"""GPU parity with radial distortion (SIMPLE_RADIAL k != 0).

The rendered scenes use pinhole cameras (k = 0), which take the pinhole
fast paths of the evaluators (`dist` = 1 exactly).  Here the same images get
a different k per view, so every tap goes through Cam::normalize's
fixed-point undistortion (the rays) and the distorted projection
`f * x * (1 + k r^2) + c` into each source (sfm-lite/src/geom.rs:27-53,
mvs/src/stereo.rs:158-228); the GPU must still match the oracle bit for bit
-- depth, cost, plane buffers and evaluation counts in every evaluator, the
fused points, and the learned schedule's plain fallback.
"""
import dataclasses

import numpy as np
import pytest

from conftest import bits, scene, view_data

pytestmark = pytest.mark.gpu

KS = [0.05, -0.04, 0.08, 0.0, -0.08, 0.02, 0.11, -0.02, 0.06, -0.06]


def _say(msg):
    print(msg, flush=True)


def _radial_views():
    views = view_data(scene("corner320"))
    return [dataclasses.replace(v, cam=dataclasses.replace(v.cam, k=KS[i % len(KS)])) for i, v in enumerate(views)]


def test_radial_every_pass(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    views = _radial_views()
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(scene("corner320").views, 4)
    opts = MvsOptions(num_sources=4, sweeps=3)
    for ref in (0, 3, 6):
        _say("radial ref %d" % ref)
        for mp in (0, 1, 6):
            dm, planes, ne = ctx.estimate_depth(ref, srcs[ref], 1.0, 10.0, opts, max_passes=mp, want_planes=True)
            d, c, n, pl, ne_o = ov.estimate_depth(ref, srcs[ref], 1.0, 10.0, opts, max_passes=mp, want_planes=True)
            assert np.array_equal(bits(planes), bits(pl)), (ref, mp)
            assert np.array_equal(bits(dm.depth), bits(d)), (ref, mp)
            assert np.array_equal(bits(dm.cost), bits(c)), (ref, mp)
            assert np.array_equal(ne[: 1 + mp], ne_o[: 1 + mp]), (ref, mp)
    # the distortion reaches the result: the pinhole views give other costs
    c0 = oracle.OracleViews(view_data(scene("corner320"))).estimate_depth(0, srcs[0], 1.0, 10.0, opts)[1]
    c1 = ov.estimate_depth(0, srcs[0], 1.0, 10.0, opts)[1]
    assert (bits(c0) != bits(c1)).mean() > 0.5


@pytest.mark.parametrize("shape", ["16x16", "16x8r", "8x8v", "16x16w", "8x8pw", "8x8px", "32x8w", "16x16sw", "24x8sw"])
def test_radial_launch_shapes(ctx, shape, monkeypatch):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    monkeypatch.setenv("MVS_TILE", shape)
    views = _radial_views()
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(scene("corner320").views, 4)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in (1, 5, 8)]
    _say("radial launch shape %s" % shape)
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    for r, s, zmin, zmax in jobs:
        d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
        assert np.array_equal(bits(out[r].depth), bits(d)), (shape, r)
        assert np.array_equal(bits(out[r].cost), bits(c)), (shape, r)


def test_radial_fusion_and_learned_fallback(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions, ProposalNet, estimate_depth_learned
    from rustcolmap_amd.scene import select_sources
    views = _radial_views()
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    sc = scene("corner320")
    srcs = select_sources(sc.views, 4)
    opts = MvsOptions(num_sources=4, sweeps=2, min_consistent=1)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    order = select_sources(sc.views, 6)
    pts = ctx.fuse_depth_maps([out[r] for r in range(len(views))], order, 0.01, opts)
    exp, _, _ = ov.fuse([out[r].depth for r in range(len(views))], order, 1, 0.03, 0.01)
    _say("radial fusion: %d points" % len(pts))
    assert len(pts) == len(exp) and len(pts) > 100
    for f in ("x", "y", "z"):
        assert np.array_equal(pts[f].view(np.uint64), exp[f].view(np.uint64)), f
    # the learned schedule with nothing accepted is the plain method, also with distortion
    net = ProposalNet(ctx, seed=2)
    sub = jobs[:3]
    lo, acc = estimate_depth_learned(ctx, net, sub, opts, vol_k=4, conf_thr=2.0)
    assert acc == 0
    for r, s, zmin, zmax in sub:
        d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
        assert np.array_equal(bits(lo[r].depth), bits(d)), r
    net.close()
