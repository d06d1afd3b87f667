"""Host-side mirror of the reference's descriptor matching over the HIP C ABI.

  PairMatches   sfm-lite/src/matching.rs:15-19
  match_pair    sfm-lite/src/matching.rs:22-98
  job_pairs     sfm-lite/src/lib.rs:197-212 (sequential window 3 or exhaustive)
  match_frames  sfm-lite/src/lib.rs:220-224, every job pair in one device batch

Descriptors are (n, 128) float32 arrays (SiftFrame.desc).  All compute runs in
libmvs_hip.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import MvsMatchStatsC, check, ptr
from .mvs import MvsContext, default_context

DEFAULT_RATIO = 0.75   # Options::default().ratio (sfm-lite/src/lib.rs:61)


@dataclass
class PairMatches:
    """Indices into frame a and frame b of each match, in increasing a."""
    a: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))

    def __len__(self):
        return len(self.a)


def _desc(d) -> np.ndarray:
    a = np.ascontiguousarray(d, np.float32)
    if a.size == 0:
        return np.zeros((0, 128), np.float32)
    if a.ndim != 2 or a.shape[1] != 128:
        raise ValueError("descriptors must be (n, 128) float32, got %r" % (a.shape,))
    return a


def match_pair(desc_a, desc_b, ratio: float = DEFAULT_RATIO, ctx: Optional[MvsContext] = None) -> PairMatches:
    """Mutual nearest neighbours of two descriptor sets that pass the
    symmetric ratio test (matching.rs:22)."""
    ctx = ctx or default_context()
    a, b = _desc(desc_a), _desc(desc_b)
    m = max(min(len(a), len(b)), 1)
    oa = np.empty(m, np.uint32)
    ob = np.empty(m, np.uint32)
    n = C.c_int64()
    check(ctx._lib.mvs_match_pair(ctx._h, ptr(a), len(a), ptr(b), len(b), float(ratio), ptr(oa), ptr(ob),
                                  C.byref(n)), "mvs_match_pair")
    return PairMatches(oa[:n.value].copy(), ob[:n.value].copy())


def job_pairs(n_frames: int, sequential: bool) -> List[Tuple[int, int]]:
    """The pair list run_sfm matches: (i, i + k) for k = 1..3 when
    sequential, else every i < j."""
    if sequential:
        return [(i, i + k) for i in range(n_frames) for k in (1, 2, 3) if i + k < n_frames]
    return [(i, j) for i in range(n_frames) for j in range(i + 1, n_frames)]


class FrameMatcher:
    """Frames' descriptors resident on one context; batches of pairs matched
    in one call."""

    def __init__(self, ctx: MvsContext, frames: Sequence[np.ndarray]):
        self.ctx = ctx
        descs = [_desc(f) for f in frames]
        self.counts = np.array([len(d) for d in descs], np.int32)
        allv = np.concatenate(descs) if descs and self.counts.sum() > 0 else np.zeros((0, 128), np.float32)
        allv = np.ascontiguousarray(allv, np.float32)
        check(ctx._lib.mvs_upload_descriptors(ctx._h, ptr(allv), ptr(self.counts), len(descs)),
              "mvs_upload_descriptors")

    def slots(self, pairs: Sequence[Tuple[int, int]]) -> np.ndarray:
        """First output slot of each pair (min(n_i, n_j) slots per pair)."""
        pr = np.asarray(pairs, np.int64).reshape(-1, 2)
        n = len(self.counts)
        off = np.zeros(len(pr) + 1, np.int64)
        if n == 0 or len(pr) == 0:
            return off
        # out-of-range indices get no slots here; the library rejects them
        ok = (pr >= 0).all(axis=1) & (pr < n).all(axis=1)
        c = self.counts.astype(np.int64)[np.where(ok[:, None], pr, 0)]
        np.cumsum(np.where(ok, c.min(axis=1), 0), out=off[1:])
        return off

    def match(self, pairs: Sequence[Tuple[int, int]], ratio: float = DEFAULT_RATIO,
              download: bool = True) -> List[PairMatches]:
        pr = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        off = self.slots(pairs)
        tot = max(int(off[-1]), 1)
        if download and (getattr(self, "_oa", None) is None or len(self._oa) < tot):
            # slot arrays kept across calls: no fresh pages to fault in per batch
            self._oa = np.empty(tot, np.uint32)
            self._ob = np.empty(tot, np.uint32)
        oa = self._oa if download else None
        ob = self._ob if download else None
        cnt = np.empty(max(len(pr), 1), np.int64)
        t0 = time.perf_counter()
        check(self.ctx._lib.mvs_match_frames(self.ctx._h, ptr(pr), len(pr), float(ratio), ptr(oa), ptr(ob),
                                             ptr(cnt)), "mvs_match_frames")
        self.last_call_ms = 1000.0 * (time.perf_counter() - t0)
        if not download:
            return [PairMatches(np.zeros(int(cnt[p]), np.uint32), np.zeros(int(cnt[p]), np.uint32))
                    for p in range(len(pr))]
        # one gather of every pair's matches out of its slots into dense
        # arrays, then a view per pair (no per-pair copies)
        m = cnt[:len(pr)]
        doff = np.zeros(len(pr) + 1, np.int64)
        np.cumsum(m, out=doff[1:])
        idx = np.repeat(off[:-1] - doff[:-1], m) + np.arange(doff[-1], dtype=np.int64)
        a, b = oa[idx], ob[idx]
        d = doff.tolist()
        return [PairMatches(a[d[p]:d[p + 1]], b[d[p]:d[p + 1]]) for p in range(len(pr))]

    def stats(self) -> dict:
        return match_stats(self.ctx)


STAT_FIELDS = ("pairs", "filtered_pairs", "distances", "matches", "overflow_items", "exact_ms", "tile_ms",
               "collect_ms", "recheck_ms", "select_ms", "host_launch_ms", "host_wait_ms", "host_out_ms")


def match_stats(ctx: MvsContext) -> dict:
    """Counts and per-stage device times (ms) of the last matching batch."""
    s = MvsMatchStatsC()
    check(ctx._lib.mvs_get_match_stats(ctx._h, C.byref(s)), "mvs_get_match_stats")
    return {k: getattr(s, k) for k in STAT_FIELDS}


def set_mode(ctx: MvsContext, exact: bool):
    """exact=True: exact f32 tiles for every pair; False (default): the
    filtered search for pairs of safe sets.  Both are bit-exact."""
    check(ctx._lib.mvs_match_set_mode(ctx._h, 1 if exact else 0), "mvs_match_set_mode")


def bound_check(ctx: MvsContext, desc_a, desc_b) -> float:
    """max |exact - MFMA estimate| / (|a|^2 + |b|^2) over all pairs of two
    sets (the filter assumes 2^-11)."""
    a, b = _desc(desc_a), _desc(desc_b)
    out = C.c_float()
    check(ctx._lib.mvs_debug_match_bound(ctx._h, ptr(a), len(a), ptr(b), len(b), C.byref(out)),
          "mvs_debug_match_bound")
    return float(out.value)


def match_frames(frames: Sequence[np.ndarray], pairs: Sequence[Tuple[int, int]], ratio: float = DEFAULT_RATIO,
                 ctx: Optional[MvsContext] = None) -> List[PairMatches]:
    """match_pair(frames[i], frames[j], ratio) for every (i, j) in `pairs`."""
    return FrameMatcher(ctx or default_context(), frames).match(pairs, ratio)


def sift_like(rng: np.random.Generator, n: int) -> np.ndarray:
    """n random SIFT-shaped descriptors: non-negative, L2-normalised,
    clipped at 0.2 and normalised again (as sift.rs makes them)."""
    d = np.abs(rng.standard_normal((n, 128))).astype(np.float32) ** 2
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    d = np.minimum(d, 0.2)
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    return d.astype(np.float32)


def synthetic_frames(n_frames: int, n_feat: int, seed: int = 0, overlap: float = 0.5,
                     noise: float = 0.03) -> List[np.ndarray]:
    """Descriptor sets of a camera sweeping past a textured world: frame f
    sees the world features [f * step, f * step + n_feat) with step =
    (1 - overlap) * n_feat, each perturbed by Gaussian noise and
    renormalised, in a random order."""
    rng = np.random.default_rng(seed)
    step = max(int(round((1.0 - overlap) * n_feat)), 1)
    world = sift_like(rng, step * (n_frames - 1) + n_feat)
    frames = []
    for f in range(n_frames):
        d = world[f * step:f * step + n_feat] + noise * rng.standard_normal((n_feat, 128)).astype(np.float32)
        d = np.maximum(d, 0.0)
        d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
        frames.append(np.ascontiguousarray(d[rng.permutation(n_feat)], np.float32))
    return frames
