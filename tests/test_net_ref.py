"""Pins of the CPU network reference (oracle/net_ref.py) used to check the GPU
kernels: fp64 autograd against central finite differences on a small input,
the parameter layout against the library's, and the loss's closed form on a
hand-made case."""
import numpy as np
import torch

from oracle import net_ref


def _data(seed, n=1, H=16, W=16):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, H, W, 64))
    x[..., :33] = rng.normal(size=(n, H, W, 33)) * 0.5
    t = np.zeros((n, H, W, 5))
    t[..., 0] = rng.uniform(0.1, 0.9, size=(n, H, W))
    nrm = rng.normal(size=(n, H, W, 3))
    nrm[..., 2] = np.abs(nrm[..., 2]) + 0.5
    t[..., 1:4] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)
    t[..., 4] = rng.random((n, H, W)) < 0.8
    return x, t


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(net_ref.num_params())
    off = 0
    for ci, co in net_ref.layer_shapes():
        p[off:off + 9 * ci * co] = rng.normal(size=9 * ci * co) * np.sqrt(2.0 / (9 * ci))
        off += 9 * ci * co + co
    p[off:off + 160] = rng.normal(size=160) * 0.2
    return p


def test_layout_matches_library_count():
    # 18 conv layers + head, as the library's Net (net_internal.h)
    assert len(net_ref.LAYERS) == 18
    assert net_ref.num_params() == 1817317


def test_autograd_matches_finite_differences_fp64():
    x, t = _data(0)
    p0 = _params(1)

    def loss_of(p):
        logits = net_ref.forward(p, x, round_acts=False, round_weights=False)
        return float(net_ref.head_loss(logits, t, 16, 16)[0])

    pt = torch.tensor(p0, requires_grad=True)
    layers_off = []
    off = 0
    for ci, co in net_ref.layer_shapes():
        layers_off.append(off)
        off += 9 * ci * co + co
    # autograd through a differentiable unpack
    logits = _forward_diff(pt, x)
    total, _ = net_ref.head_loss(logits, t, 16, 16)
    total.backward()
    g = pt.grad.numpy()
    rng = np.random.default_rng(3)
    picks = [layers_off[0] + 5, layers_off[3] + 100, layers_off[9] + 7, layers_off[10] + 1234, layers_off[17] + 3,
             off + 10, off + 161]
    picks += list(rng.integers(0, net_ref.num_params(), size=5))
    for i in picks:
        e = np.zeros_like(p0)
        e[i] = 1e-6
        fd = (loss_of(p0 + e) - loss_of(p0 - e)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def _forward_diff(pt, x):
    import torch.nn.functional as F
    off = 0
    ws = []
    for ci, co in net_ref.layer_shapes():
        w = pt[off:off + 9 * ci * co].reshape(3, 3, ci, co).permute(3, 2, 0, 1)
        off += 9 * ci * co
        ws.append((w, pt[off:off + co]))
        off += co
    hw = pt[off:off + 160].reshape(5, 32)
    hb = pt[off + 160:off + 165]
    acts = {"IN": torch.as_tensor(x, dtype=torch.float64).permute(0, 3, 1, 2)}
    for l, (x0, x1, o) in enumerate(net_ref.LAYERS):
        if x0 in net_ref.POOLS:
            acts[x0] = F.avg_pool2d(acts[net_ref.POOLS[x0]], 2)
        inp = acts[x0]
        if x1:
            inp = torch.cat([inp, F.interpolate(acts[x1], scale_factor=2, mode="nearest")], 1)
        acts[o] = F.relu(F.conv2d(inp, ws[l][0], ws[l][1], padding=1))
    return acts["D0"].permute(0, 2, 3, 1) @ hw.T + hb


def test_head_loss_closed_form():
    logits = torch.zeros((1, 1, 2, 5), dtype=torch.float64)
    logits[0, 0, 0] = torch.tensor([0.0, 0.0, 0.0, 2.0, 0.0])     # inv 0.5, normal (0,0,1), conf 0.5
    logits[0, 0, 1] = torch.tensor([0.0, 0.0, 3.0, 4.0, 0.0])     # normal (0, .6, .8)
    t = np.zeros((1, 1, 2, 5))
    t[0, 0, 0] = [0.25, 0, 0, 1, 1]
    t[0, 0, 1] = [0.5, 0, 0, 1, 1]
    total, (ld, ln, lc, nv) = net_ref.head_loss(logits, t, 1, 2)
    assert abs(float(ld) - 0.25) < 1e-12
    assert abs(float(ln) - ((1 - 2 / (2 + 1e-6)) + (1 - 4 / (5 + 1e-6)))) < 1e-12
    assert abs(float(lc) - 2 * np.log(2)) < 1e-12 and float(nv) == 2
    assert abs(float(total) - 0.5 * (0.25 + 0.5 * float(ln) + 0.25 * 2 * np.log(2))) < 1e-12
    # one valid pixel of two: depth and normal are means over valid pixels,
    # the confidence term a mean over all pixels
    t[0, 0, 1, 4] = 0.0
    total, (ld, ln, lc, nv) = net_ref.head_loss(logits, t, 1, 2)
    ln0 = 1 - 2 / (2 + 1e-6)
    assert float(nv) == 1 and abs(float(ld) - 0.25) < 1e-12 and abs(float(ln) - ln0) < 1e-12
    assert abs(float(total) - (0.25 + 0.5 * ln0 + 0.25 * 2 * np.log(2) / 2)) < 1e-12
    # the confidence target is 1 only at a valid target within 2 %: an
    # invalid pixel and a valid one 3 % off are negatives
    logits[0, 0, 0, 4] = 1.0                                          # pixel 0: conf = sigmoid(1)
    logits[0, 0, 1, 4] = 2.0                                          # pixel 1 (invalid): sigmoid(2)
    s1, s2 = 1 / (1 + np.exp(-1.0)), 1 / (1 + np.exp(-2.0))
    t[0, 0, 0] = [0.5 / 1.03, 0, 0, 1, 1]                             # inv 0.5 is 3 % above the target
    total, (ld, ln, lc, nv) = net_ref.head_loss(logits, t, 1, 2)
    assert abs(float(lc) - (-np.log(1 - s1) - np.log(1 - s2))) < 1e-12
    t[0, 0, 0] = [0.5 / 1.01, 0, 0, 1, 1]                             # within 2 %: positive
    total, (ld, ln, lc, nv) = net_ref.head_loss(logits, t, 1, 2)
    assert abs(float(lc) - (-np.log(s1) - np.log(1 - s2))) < 1e-12
