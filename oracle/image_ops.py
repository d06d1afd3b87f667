"""CPU restatement of the input stage of `load_views` (mvs/src/lib.rs:185-197)
-- TEST INFRASTRUCTURE ONLY: tests use it as the checker of the library's
input-stage kernels; the library never imports it.

The reference resizes each decoded image with the third-party `image` crate,
version 0.24.9 (mvs/Cargo.toml; not vendored in the reference tree), as
`img.resize_exact(mw, mh, FilterType::Triangle)` and then converts with
`to_luma8()` / `to_rgb8()`.  Restated here from that crate's published
algorithm (imageops/sample.rs `resize`, `vertical_sample`,
`horizontal_sample`; color.rs `rgb_to_luma`):

* resize: a vertical pass into f32, then a horizontal pass; for each output
  row (column) with ratio = in / out (f32), sratio = max(ratio, 1),
  support = 1 * sratio, centre = (o + 0.5) * ratio, the taps are
  left = clamp(floor(centre - support), 0, in - 1) ..
  right = clamp(ceil(centre + support), left + 1, in), weights
  triangle((i - (centre - 0.5)) / sratio) = max(1 - |x|, 0) normalised by
  their f32 sum (accumulated in tap order), and the output is
  sum over taps of value * weight accumulated in f32 in tap order;
  the horizontal pass clamps to [0, 255] and rounds half away from zero.
* luma of RGB (color.rs `rgb_to_luma` with `SRGB_LUMA` = [2126, 7152, 722]
  and `SRGB_LUMA_DIV` = 10000): the weighted sum in integers of the
  `Enlargeable` larger type (u32 for u8), divided by 10000 with truncation;
  a one-channel image is its own luma.

Parity: the crate cannot be run here and the reference ships no vectors for
this stage.  The resize restatement is pinned by closed forms only (exact
2x downscale weights [1, 3, 3, 1] / 8, constant images stay constant,
identity at equal size, separability) and stays "parity unpinned" for
general inputs.  The luma is integer arithmetic whose closed forms fix it:
every grey (v, v, v) maps to v (the weights sum to the divisor), the
primaries to floor(255 w / 10000) and each value to floor(sum / 10000);
see tests/test_input_stage.py.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _taps(n_in: int, n_out: int):
    """Per output index: (left, weights f32 array) -- sample.rs's window."""
    ratio = F(n_in) / F(n_out)
    sratio = ratio if ratio >= F(1.0) else F(1.0)
    support = F(1.0) * sratio
    out = []
    for o in range(n_out):
        centre = (F(o) + F(0.5)) * ratio
        left = int(np.floor(centre - support))
        left = min(max(left, 0), n_in - 1)
        right = int(np.ceil(centre + support))
        right = min(max(right, left + 1), n_in)
        c = centre - F(0.5)
        ws = []
        s = F(0.0)
        for i in range(left, right):
            x = (F(i) - c) / sratio
            ax = abs(x)
            w = F(1.0) - ax if ax < F(1.0) else F(0.0)
            ws.append(F(w))
            s = F(s + F(w))
        out.append((left, np.array([F(w / s) for w in ws], np.float32)))
    return out


def resize_triangle(img: np.ndarray, mw: int, mh: int) -> np.ndarray:
    """u8 (h, w) or (h, w, c) -> u8 (mh, mw[, c])."""
    a = np.asarray(img, np.uint8)
    squeeze = a.ndim == 2
    if squeeze:
        a = a[:, :, None]
    h, w, c = a.shape
    if (mw, mh) == (w, h):
        return a[:, :, 0].copy() if squeeze else a.copy()
    src = a.astype(np.float32)
    tmp = np.zeros((mh, w, c), np.float32)
    for oy, (left, ws) in enumerate(_taps(h, mh)):
        t = np.zeros((w, c), np.float32)
        for k, wt in enumerate(ws):
            t = (t + src[left + k] * wt).astype(np.float32)
        tmp[oy] = t
    out = np.zeros((mh, mw, c), np.float32)
    for ox, (left, ws) in enumerate(_taps(w, mw)):
        t = np.zeros((mh, c), np.float32)
        for k, wt in enumerate(ws):
            t = (t + tmp[:, left + k] * wt).astype(np.float32)
        out[:, ox] = t
    out = np.clip(out, F(0.0), F(255.0))
    # f32::round (half away from zero) of a non-negative value, from the exact
    # fraction (x + 0.5 could itself round up)
    frac = out - np.floor(out)
    r = np.where(frac < F(0.5), np.floor(out), np.floor(out) + 1)
    res = r.astype(np.uint8)
    return res[:, :, 0] if squeeze else res


SRGB_LUMA = (2126, 7152, 722)
SRGB_LUMA_DIV = 10000


def rgb_to_luma(rgb: np.ndarray) -> np.ndarray:
    """u8 (h, w, 3) -> u8 (h, w): (2126 r + 7152 g + 722 b) / 10000 in u32,
    truncating (the result never exceeds 255, so the clamp is a no-op)."""
    p = np.asarray(rgb, np.uint8).astype(np.uint32)
    l = SRGB_LUMA[0] * p[..., 0] + SRGB_LUMA[1] * p[..., 1] + SRGB_LUMA[2] * p[..., 2]
    return (l // SRGB_LUMA_DIV).astype(np.uint8)


def load_view_pixels(img: np.ndarray, mw: int, mh: int):
    """(gray, rgb) at the working size as load_views makes them from a
    decoded u8 image (gray or RGB)."""
    small = resize_triangle(img, mw, mh)
    if small.ndim == 2:
        return small, np.repeat(small[:, :, None], 3, axis=2)
    return rgb_to_luma(small), small
