"""The pass kernels' current-plane repeat rule changes no result (GPU).

A neighbour candidate equal to the current plane at the d_plane the plane
was costed with is counted without an evaluation (JobDesc.costs_exact).
The oracle parity tests cover it for the planes the init sweep and the
passes store; the learned initialisation also stores proposal planes, slanted
ones, whose d_plane its proposal-cost kernel records (JobDesc.dplane).  Here
every result -- depth, cost and normals, and the pass evaluation counts -- is
compared bit for bit with the same runs with the rule switched off
(MVS_COSTS_EXACT=0; the first-neighbour rule needs no condition and stays on),
for the learned initialisation in both of its modes and for the plain batch.
"""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _scene():
    from rustcolmap_amd.learned import training_spec
    from rustcolmap_amd.mvs import ViewData
    from rustcolmap_amd.scene import render_scene, select_sources
    sc = render_scene(training_spec(5001), with_gt=False)
    views = [ViewData.from_scene_view(v) for v in sc.views]
    srcs = select_sources(sc.views, 4)
    return views, [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]


def _same(a, b, tag):
    for r in a:
        assert np.array_equal(bits(a[r].depth), bits(b[r].depth)), (tag, r)
        assert np.array_equal(bits(a[r].cost), bits(b[r].cost)), (tag, r)
        assert np.array_equal(bits(a[r].normals), bits(b[r].normals)), (tag, r)


@pytest.mark.parametrize("extra", [True, False])
def test_learned_runs_do_not_depend_on_the_current_plane_rule(ctx, monkeypatch, extra):
    from rustcolmap_amd.learned import DEFAULT_WEIGHTS, load_weights
    from rustcolmap_amd.mvs import MvsOptions, ProposalNet, estimate_depth_learned
    views, jobs = _scene()
    ctx.upload_views(views)
    net = ProposalNet(ctx, seed=0)
    load_weights(net, DEFAULT_WEIGHTS)
    opts = MvsOptions(num_sources=4, sweeps=3)
    _say("learned, extra=%s, rule on" % extra)
    on, acc_on = estimate_depth_learned(ctx, net, jobs, opts, vol_k=4, conf_thr=0.3, extra=extra)
    ne_on = ctx.stats().n_eval_pass
    monkeypatch.setenv("MVS_COSTS_EXACT", "0")
    _say("learned, extra=%s, rule off" % extra)
    off, acc_off = estimate_depth_learned(ctx, net, jobs, opts, vol_k=4, conf_thr=0.3, extra=extra)
    ne_off = ctx.stats().n_eval_pass
    assert acc_on == acc_off and acc_on > 0          # proposals were accepted: slanted planes are present
    _same(on, off, "learned extra=%s" % extra)
    assert ne_on == ne_off


@pytest.mark.parametrize("shape", ["16x16sw", "8x8pw"])
def test_plain_batches_do_not_depend_on_the_current_plane_rule(ctx, monkeypatch, shape):
    from rustcolmap_amd.mvs import MvsOptions
    views, jobs = _scene()
    ctx.upload_views(views)
    monkeypatch.setenv("MVS_TILE", shape)
    opts = MvsOptions(num_sources=4, sweeps=4)
    _say("plain %s, rule on" % shape)
    on = ctx.estimate_depth_batch(jobs, opts, mode=0)
    ne_on = ctx.stats().n_eval_pass
    monkeypatch.setenv("MVS_COSTS_EXACT", "0")
    _say("plain %s, rule off" % shape)
    off = ctx.estimate_depth_batch(jobs, opts, mode=0)
    ne_off = ctx.stats().n_eval_pass
    _same(on, off, "plain " + shape)
    assert ne_on == ne_off
