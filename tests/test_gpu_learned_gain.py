"""What the trained proposal network gains over the method without it.

Held-out scenes of the training distribution (seeds 5000-5002, never used
for training), 8 views of 640x480, K = 4.  The reference run is the plain method
with 8 sweeps; both challengers get 2 sweeps: the plain method, and the
learned initialisation with the committed weights
(rustcolmap_amd/weights/proposal_unet.npz) at confidence 0.5, where an
accepted proposal joins its pixel's plain sweep as a 49th hypothesis.  A
network with random weights stands in for "no training" as a third
reference.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _agree(ref, other, tol=0.01):
    m = ref > 0
    o = other[m]
    ok = (o > 0) & (np.abs(1.0 / np.where(o > 0, o, 1.0) - 1.0 / ref[m]) <= tol / ref[m])
    return float(ok.mean()) if m.any() else float("nan")


def test_trained_network_gains_over_plain_two_sweeps(ctx):
    from rustcolmap_amd.learned import DEFAULT_WEIGHTS, load_weights, training_spec
    from rustcolmap_amd.mvs import MvsOptions, ProposalNet, ViewData, estimate_depth_learned
    from rustcolmap_amd.scene import render_scene, select_sources
    trained = ProposalNet(ctx, seed=0)
    load_weights(trained, DEFAULT_WEIGHTS)
    untrained = ProposalNet(ctx, seed=0)
    stats = {"plain2": [], "learned2": [], "untrained2": []}
    evals = {"plain2": 0, "learned2": 0, "untrained2": 0}
    acc_total = 0
    npix = 0
    for seed in (5000, 5001, 5002):
        sc = render_scene(training_spec(seed), with_gt=False)
        views = [ViewData.from_scene_view(v) for v in sc.views]
        ctx.upload_views(views)
        srcs = select_sources(sc.views, 4)
        jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
        ref8 = ctx.estimate_depth_batch(jobs, MvsOptions(num_sources=4, sweeps=8), mode=4)
        short = MvsOptions(num_sources=4, sweeps=2)
        plain2 = ctx.estimate_depth_batch(jobs, short, mode=4)
        evals["plain2"] += ctx.stats().n_eval_init
        learned2, acc = estimate_depth_learned(ctx, trained, jobs, short, vol_k=4, conf_thr=0.5)
        evals["learned2"] += ctx.stats().n_eval_init
        acc_total += acc
        untrained2, _ = estimate_depth_learned(ctx, untrained, jobs, short, vol_k=4, conf_thr=0.5)
        evals["untrained2"] += ctx.stats().n_eval_init
        npix += sum(v.w * v.h for v in views)
        for name, out in (("plain2", plain2), ("learned2", learned2), ("untrained2", untrained2)):
            for r in range(len(views)):
                stats[name].append((_agree(ref8[r].depth, out[r].depth), float((out[r].depth > 0).mean()),
                                    float((ref8[r].depth > 0).mean())))
    summary = {k: (float(np.nanmean([a for a, _, _ in v])), float(np.mean([b for _, b, _ in v]))) for k, v in
               stats.items()}
    _say("agreement with plain-8 / valid fraction: %s; 8-sweep valid %.4f; accepted %.4f; init evaluations %s"
         % (summary, float(np.mean([c for _, _, c in stats["plain2"]])), acc_total / npix, evals))
    assert acc_total > 0
    # every pixel ran the plain sweep, the accepted ones one proposal more ...
    assert evals["learned2"] == evals["plain2"] + acc_total
    # ... and after two sweeps the result agrees with the 8-sweep run more
    # often than the plain method with two sweeps and than with an untrained
    # network
    assert summary["learned2"][0] > summary["plain2"][0]
    assert summary["learned2"][0] > summary["untrained2"][0]
    trained.close()
    untrained.close()
