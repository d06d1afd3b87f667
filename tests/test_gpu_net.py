"""GPU tests of the learned plane-proposal path.

* the 32-plane sweep volume equals the oracle plane_cost's exact window
  census cost (Hamming sum + 48 per invalid tap, no abort) bit for bit, at
  full resolution and on the stride-2 grid;
* the BF16/MFMA forward matches the CPU reference (oracle/net_ref.py) run
  with the same bf16 roundings, to bf16 tolerance;
* the GPU gradients of every parameter tensor match fp64 CPU autograd
  (tolerance: relative norm error 5 %, cosine 0.995 per tensor);
* Adam training on seeded scenes lowers the loss;
* with a confidence threshold no pixel reaches, the learned schedule
  reproduces the plain method bit for bit (the fallback is the plain init),
  in both modes (proposal replaces / proposal as an extra hypothesis); with
  every proposal accepted both still produce valid, gated maps.
"""
import numpy as np
import pytest

from conftest import bits, scene, view_data

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_sweep_volume_is_window_census_cost(ctx):
    """Each volume entry is the plane's census cost over the 5x5 window and
    the sources summed exactly -- the oracle plane_cost's Hamming sum plus 48
    per invalid tap, evaluated with no abort -- divided once by 48 * 25 * K."""
    import oracle
    from rustcolmap_amd.mvs import sweep_volume
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(sc.views, 4)
    ref = 5
    vol = sweep_volume(ctx, (ref, srcs[ref], 1.0, 10.0))
    vol2 = sweep_volume(ctx, (ref, srcs[ref], 1.0, 10.0), stride=2)
    assert vol.shape == (240, 320, 32) and vol2.shape == (120, 160, 32)
    # the coarse grid samples the full-resolution volume
    assert np.array_equal(bits(vol2), bits(vol[::2, ::2]))
    ws = ov.workspace(ref, srcs[ref], 1.0, 10.0)
    inv_min, inv_max = np.float32(1.0) / np.float32(10.0), np.float32(1.0) / np.float32(1.0)
    rng = np.random.default_rng(0)
    for (x, y) in [(0, 0), (319, 239), (1, 238)] + [tuple(p) for p in rng.integers(0, [320, 240], size=(40, 2))]:
        for h in range(32):
            inv = np.float32(inv_min + np.float32(np.float32(inv_max - inv_min) * np.float32(h)) / np.float32(31))
            d = np.float32(np.float32(1.0) / inv)
            _, info = ws.plane_cost(x, y, 0.0, 0.0, 1.0, d, np.inf, info=True)
            want = np.float32(info.ham_sum + 48 * info.n_invalid) / np.float32(48 * 25 * 4)
            assert np.float32(vol[y, x, h]).view(np.uint32) == want.view(np.uint32), (x, y, h)


def _rand_batch(rng, n, Hp, Wp, h, w):
    x = np.zeros((n, Hp, Wp, 64), np.float32)
    x[:, :h, :w, :33] = rng.normal(size=(n, h, w, 33)).astype(np.float32) * 0.5
    t = np.zeros((n, h, w, 5), np.float32)
    t[..., 0] = rng.uniform(0.05, 0.95, size=(n, h, w))
    nrm = rng.normal(size=(n, h, w, 3))
    nrm[..., 2] = np.abs(nrm[..., 2]) + 0.5
    t[..., 1:4] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    t[..., 4] = rng.random((n, h, w)) < 0.7
    return x, t


def test_forward_matches_cpu_reference(ctx):
    import torch
    from oracle import net_ref
    from rustcolmap_amd.mvs import ProposalNet
    net = ProposalNet(ctx, seed=3)
    params = net.get_params()
    assert params.size == net_ref.num_params()
    rng = np.random.default_rng(1)
    x, t = _rand_batch(rng, 2, 32, 48, 30, 45)
    _, pred, _ = net.debug_grads(x, t, 30, 45)
    logits = net_ref.forward(params, x, round_acts=True).numpy()[:, :30, :45]
    inv = 1 / (1 + np.exp(-logits[..., 0]))
    conf = 1 / (1 + np.exp(-logits[..., 4]))
    nv = logits[..., 1:4]
    n = nv / (np.linalg.norm(nv, axis=-1, keepdims=True) + 1e-6)
    _say("forward max err inv %.3g conf %.3g normal %.3g" % (np.abs(pred[..., 0] - inv).max(),
                                                           np.abs(pred[..., 4] - conf).max(),
                                                           np.abs(pred[..., 1:4] - n).max()))
    assert np.abs(pred[..., 0] - inv).max() < 2e-2
    assert np.abs(pred[..., 4] - conf).max() < 2e-2
    assert np.mean(np.abs(pred[..., 1:4] - n)) < 1e-2
    net.close()
    del torch


def test_gradients_match_cpu_autograd(ctx):
    from oracle import net_ref
    from rustcolmap_amd.mvs import ProposalNet
    net = ProposalNet(ctx, seed=5)
    params = net.get_params()
    rng = np.random.default_rng(2)
    x, t = _rand_batch(rng, 2, 32, 32, 32, 32)
    g_gpu, _, loss = net.debug_grads(x, t, 32, 32)
    g_ref, _, parts = net_ref.grads(params, x, t, 32, 32, round_acts=True)
    off = 0
    shapes = net_ref.layer_shapes()
    for l, (ci, co) in enumerate(shapes):
        for name, n in (("w", 9 * ci * co), ("b", co)):
            a, b = g_gpu[off:off + n].astype(np.float64), g_ref[off:off + n]
            if l == 0 and name == "w":
                # input channels 33..63 are zero: their gradients vanish on both sides
                a = a.reshape(9, ci, co)[:, :33].ravel()
                b = b.reshape(9, ci, co)[:, :33].ravel()
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            _say("layer %2d %s rel %.4f cos %.5f" % (l, name, rel, cos))
            assert rel < 0.05 and cos > 0.995, (l, name, rel, cos)
            off += n
    a, b = g_gpu[off:].astype(np.float64), g_ref[off:]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 0.05
    assert abs(loss[3] - parts[3]) < 0.5            # valid pixel count
    net.close()


def test_adam_training_lowers_loss(ctx):
    from rustcolmap_amd.learned import build_training_set, train
    from rustcolmap_amd.mvs import ProposalNet
    net = ProposalNet(ctx, seed=7)
    n = build_training_set(ctx, net, n_scenes=3, size=(160, 128), views_per_scene=4, seed0=9000, log=_say)
    assert n >= 8
    everything = np.arange(n)

    def objective():
        # a step with lr = 0 leaves the weights as they are and reports the
        # loss terms of the whole set: depth L1 + 0.5 normal + 0.25 BCE
        l = net.train_step(everything, 0.0)
        return float(l[0] + 0.5 * l[1] + 0.25 * l[2]), l
    before, l0 = objective()
    train(net, n, steps=60, batch=4, lr=1e-3, seed=0, log=_say)
    after, l1 = objective()
    _say("objective on all %d samples: %.4f -> %.4f (terms %s -> %s)" % (n, before, after, l0, l1))
    assert after < 0.9 * before
    net.close()


def test_learned_schedule_fallback_is_plain_and_accept_all_runs(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions, ProposalNet, estimate_depth_learned
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(sc.views, 4)
    opts = MvsOptions(num_sources=4, sweeps=2)
    net = ProposalNet(ctx, seed=1)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in (1, 5, 8)]
    out, acc = estimate_depth_learned(ctx, net, jobs, opts, vol_k=4, conf_thr=2.0)
    assert acc == 0
    for r, s, zmin, zmax in jobs:
        d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
        assert np.array_equal(bits(out[r].depth), bits(d)), r
        assert np.array_equal(bits(out[r].cost), bits(c)), r
    # extra-hypothesis mode: with nothing accepted it is the plain method too
    out, acc = estimate_depth_learned(ctx, net, jobs, opts, vol_k=4, conf_thr=2.0, extra=True)
    assert acc == 0
    for r, s, zmin, zmax in jobs:
        d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
        assert np.array_equal(bits(out[r].depth), bits(d)), r
        assert np.array_equal(bits(out[r].cost), bits(c)), r
    for extra in (False, True):
        out, acc = estimate_depth_learned(ctx, net, jobs, opts, vol_k=4, conf_thr=0.0, extra=extra)
        _say("accepted %d proposals (extra=%s)" % (acc, extra))
        assert acc > 0
        for r, s, zmin, zmax in jobs:
            z = out[r].depth
            assert np.all(np.isfinite(z)) and np.all((z == 0) | ((z > 0.5) & (z < 20.0)))
    net.close()


def test_network_outlives_its_context():
    """Destroying a context first orphans its networks: their calls fail
    with a state error and destroying them afterwards is safe."""
    from rustcolmap_amd import _lib
    from rustcolmap_amd.mvs import MvsContext, ProposalNet
    c2 = MvsContext(0)
    net = ProposalNet(c2, seed=2)
    assert net.get_params().size == net.num_params
    c2.close()
    with pytest.raises(_lib.MvsError):
        net.get_params()
    net.close()
