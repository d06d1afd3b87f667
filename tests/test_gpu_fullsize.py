"""Parity at the real shape: 20 views x 1920x1080, K = 8, 8 sweeps.

The oracle cannot run whole 1080p views in test time, so it computes row
bands: a pass reads only the rows next to each pixel, so after 16 passes a
band computed with a 16-row margin is exact in its interior.  Those rows are
compared bit for bit with the GPU's full-view result; the rest of every view
is checked through size-independent properties (run-to-run identity, the
gate's closed form, unit normals, depth range).
"""
import numpy as np
import pytest

from conftest import bits, view_data

pytestmark = pytest.mark.gpu

_FULL = {}


def _say(msg):
    print(msg, flush=True)


def _full_scene():
    if "sc" not in _FULL:
        from rustcolmap_amd.scene import config_spec, render_scene
        _say("rendering room1080 (20 x 1920x1080)")
        _FULL["sc"] = render_scene(config_spec("room1080"), device="cuda", with_gt=True)
    return _FULL["sc"]


def test_fullsize_bands_bitexact_and_properties(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = _full_scene()
    views = view_data(sc)
    K, sweeps = 8, 8
    srcs = select_sources(sc.views, K)
    opts = MvsOptions(num_sources=K, sweeps=sweeps)
    ctx.upload_views(views)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    _say("gpu batch 1")
    out1 = ctx.estimate_depth_batch(jobs, opts, mode=0)
    _say("gpu batch 2")
    out2 = ctx.estimate_depth_batch(jobs, opts, mode=0)
    for r in range(len(views)):
        assert np.array_equal(bits(out1[r].depth), bits(out2[r].depth)), r
        assert np.array_equal(bits(out1[r].cost), bits(out2[r].cost)), r
    ov = oracle.OracleViews(views)
    P = 2 * sweeps
    for r, y0 in ((0, 0), (10, 530), (19, 1080 - 12)):
        lo, hi = max(0, y0 - P), min(1080, y0 + 12 + P)
        _say("oracle band view %d rows %d-%d" % (r, lo, hi))
        d, c, n, _, _ = ov.estimate_depth(r, srcs[r], 1.0, 10.0, opts, rows=(lo, hi))
        a, b = y0, min(1080, y0 + 12)
        assert np.array_equal(bits(out1[r].depth[a:b]), bits(d[a:b])), (r, a, b)
        assert np.array_equal(bits(out1[r].cost[a:b]), bits(c[a:b])), (r, a, b)
        assert np.array_equal(bits(out1[r].normals[a:b]), bits(n[a:b])), (r, a, b)
    # properties over every pixel of every view
    for r in range(len(views)):
        cen = ctx.view_census(r)
        tex = np.zeros(cen.shape, np.int32)
        x = cen.copy()
        while np.any(x):
            tex += (x & np.uint64(1)).astype(np.int32)
            x >>= np.uint64(1)
        dm = out1[r]
        z = dm.depth
        valid = z > 0
        assert np.all(valid <= ((tex >= 4) & (tex <= 44) & (dm.cost < np.float32(0.38))))
        assert np.all((z[valid] > 0.5) & (z[valid] < 20.0))
        nn = np.linalg.norm(dm.normals.astype(np.float64), axis=-1)
        assert np.all(np.abs(nn - 1.0) < 1e-5)
        assert np.all((dm.cost >= 0) & (dm.cost <= 1.0))


def test_fullsize_fusion_matches_oracle(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = _full_scene()
    views = view_data(sc)[:6]
    ctx.upload_views(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(sc.views[:6], 4)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(6)]
    _say("gpu batch for fusion")
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    order = select_sources(sc.views[:6], 5)
    _say("gpu fusion")
    got = ctx.fuse_resident([True] * 6, order, 0.005, opts)
    ov = oracle.OracleViews(views)
    _say("oracle fusion")
    exp, ncand, _ = ov.fuse([out[r].depth for r in range(6)], order, 2, 0.03, 0.005)
    assert ctx.stats().n_candidates == ncand
    assert len(got) == len(exp) and len(got) > 1000
    for f in ("x", "y", "z"):
        assert np.array_equal(bits(got[f]), bits(exp[f])), f
    assert np.array_equal(got["rgb"], exp["rgb"])
