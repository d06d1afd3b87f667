"""TEST INFRASTRUCTURE ONLY -- CPU reference of the plane-proposal network.

The network has no counterpart in the reference repository; this module is
the numerics reference that the HIP forward, backward and Adam kernels are
checked against: the same 5-level U-Net written with torch operations on the
CPU (fp32 or fp64), parameters taken from the library's flat fp32 vector.

Parameter layout (per layer l = 0..17): weights [9][cin][cout] with tap
t = 3 * ky + kx, then bias [cout]; then the head weights [5][32] and bias [5].
Activations optionally rounded to bf16 after every convolution and pooling,
which is what the GPU stores.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

ACT_CH = {"IN": 64, "A0": 32, "E0": 32, "P0": 32, "A1": 64, "E1": 64, "P1": 64, "A2": 96, "E2": 96, "P2": 96,
          "A3": 128, "E3": 128, "P3": 128, "A4": 160, "E4": 160, "B3": 128, "D3": 128, "B2": 96, "D2": 96,
          "B1": 64, "D1": 64, "B0": 32, "D0": 32}
LAYERS = [("IN", None, "A0"), ("A0", None, "E0"), ("P0", None, "A1"), ("A1", None, "E1"), ("P1", None, "A2"),
          ("A2", None, "E2"), ("P2", None, "A3"), ("A3", None, "E3"), ("P3", None, "A4"), ("A4", None, "E4"),
          ("E3", "E4", "B3"), ("B3", None, "D3"), ("E2", "D3", "B2"), ("B2", None, "D2"), ("E1", "D2", "B1"),
          ("B1", None, "D1"), ("E0", "D1", "B0"), ("B0", None, "D0")]
POOLS = {"P0": "E0", "P1": "E1", "P2": "E2", "P3": "E3"}


def layer_shapes():
    out = []
    for x0, x1, o in LAYERS:
        cin = ACT_CH[x0] + (ACT_CH[x1] if x1 else 0)
        out.append((cin, ACT_CH[o]))
    return out


def num_params() -> int:
    return sum(9 * ci * co + co for ci, co in layer_shapes()) + 5 * 32 + 5


def unpack(params: np.ndarray, dtype=torch.float64):
    """Flat vector -> list of (weight [cout][cin][3][3], bias) + head (W [5][32], b [5])."""
    p = torch.as_tensor(np.asarray(params), dtype=dtype)
    off = 0
    layers = []
    for ci, co in layer_shapes():
        w = p[off:off + 9 * ci * co].reshape(3, 3, ci, co).permute(3, 2, 0, 1).contiguous()
        off += 9 * ci * co
        b = p[off:off + co]
        off += co
        layers.append((w, b))
    hw = p[off:off + 160].reshape(5, 32)
    hb = p[off + 160:off + 165]
    return layers, (hw, hb)


def bf16(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.bfloat16).to(t.dtype)


def forward(params, x_nhwc, round_acts=True, round_weights=True, dtype=torch.float64):
    """x_nhwc: (N, H, W, 64) input. Returns the head logits (N, H, W, 5)."""
    layers, (hw, hb) = unpack(params, dtype)
    if round_weights:
        layers = [(bf16(w), b) for w, b in layers]
    x = torch.as_tensor(np.asarray(x_nhwc), dtype=dtype).permute(0, 3, 1, 2)
    if round_weights:
        x = bf16(x)
    acts = {"IN": x}
    rnd = bf16 if round_acts else (lambda t: t)
    for l, (x0, x1, o) in enumerate(LAYERS):
        if x0 in POOLS:
            acts[x0] = rnd(F.avg_pool2d(acts[POOLS[x0]], 2))
        inp = acts[x0]
        if x1:
            inp = torch.cat([inp, F.interpolate(acts[x1], scale_factor=2, mode="nearest")], 1)
        w, b = layers[l]
        acts[o] = rnd(F.relu(F.conv2d(inp, w, b, padding=1)))
    d0 = acts["D0"].permute(0, 2, 3, 1)
    return d0 @ hw.T + hb


def head_loss(logits, target, h, w, w_depth=1.0, w_normal=0.5, w_conf=0.25, conf_tol=0.02):
    """Loss of the library's head on the h x w image inside the padded logits.
    target: (N, h, w, 5) = inverse depth (normalised), normal (3), validity.
    Depth L1 and normal (1 - cos) are means over the valid pixels, the
    confidence BCE a mean over all pixels, against "within conf_tol of a
    valid target"."""
    o = logits[:, :h, :w, :]
    t = torch.as_tensor(np.asarray(target), dtype=o.dtype)
    inv = torch.sigmoid(o[..., 0])
    nv = o[..., 1:4]
    n = nv / (torch.linalg.norm(nv, dim=-1, keepdim=True) + 1e-6)
    conf = torch.sigmoid(o[..., 4])
    v = (t[..., 4] > 0.5).to(o.dtype)
    err = inv - t[..., 0]
    good = ((v > 0) & (err.abs().detach() <= conf_tol * torch.clamp(t[..., 0], min=1e-3))).to(o.dtype)
    ld = (v * err.abs()).sum()
    ln = (v * (1.0 - (n * t[..., 1:4]).sum(-1))).sum()
    c = torch.clamp(conf, 1e-7, 1 - 1e-7)
    lc = -(good * torch.log(c) + (1 - good) * torch.log(1 - c)).sum()
    nv = max(float(v.sum()), 1.0)
    total = (w_depth * ld + w_normal * ln) / nv + w_conf * lc / (o.shape[0] * h * w)
    return total, (ld, ln, lc, v.sum())


def grads(params, x_nhwc, target, h, w, round_acts=False):
    """fp64 autograd gradient of the loss w.r.t. the flat parameter vector
    (weights taken bf16-rounded, as the GPU uses them)."""
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    layers, (hw, hb) = unpack(p.detach().numpy())
    # rebuild with a differentiable view of p
    off = 0
    ws = []
    for ci, co in layer_shapes():
        wv = p[off:off + 9 * ci * co].reshape(3, 3, ci, co).permute(3, 2, 0, 1)
        wq = wv + (bf16(wv.detach()) - wv.detach())      # bf16 value, identity gradient
        off += 9 * ci * co
        ws.append((wq, p[off:off + co]))
        off += co
    hw = p[off:off + 160].reshape(5, 32)
    hb = p[off + 160:off + 165]
    x = bf16(torch.as_tensor(np.asarray(x_nhwc), dtype=torch.float64)).permute(0, 3, 1, 2)
    acts = {"IN": x}
    rnd = (lambda t: t + (bf16(t.detach()) - t.detach())) if round_acts else (lambda t: t)
    for l, (x0, x1, o) in enumerate(LAYERS):
        if x0 in POOLS:
            acts[x0] = rnd(F.avg_pool2d(acts[POOLS[x0]], 2))
        inp = acts[x0]
        if x1:
            inp = torch.cat([inp, F.interpolate(acts[x1], scale_factor=2, mode="nearest")], 1)
        acts[o] = rnd(F.relu(F.conv2d(inp, ws[l][0], ws[l][1], padding=1)))
    logits = acts["D0"].permute(0, 2, 3, 1) @ hw.T + hb
    total, parts = head_loss(logits, target, h, w)
    total.backward()
    return p.grad.numpy(), logits.detach().numpy(), [float(v) for v in parts]
