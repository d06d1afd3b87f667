"""Descriptor matching on the device against the oracle
(sfm-lite/src/matching.rs:22-98, lib.rs:197-224), bit-exact: the same match
list, in the same order -- through both searches (exact f32 tiles, and the
bf16x3 MFMA filter with exact recheck), plus the filter's error bound
measured against its assumption."""
import numpy as np
import pytest

import oracle
from rustcolmap_amd.matching import (FrameMatcher, bound_check, job_pairs, match_pair, match_stats, set_mode,
                                     sift_like, synthetic_frames)

pytestmark = pytest.mark.gpu

EPS_REL = 2.0 ** -11   # the filter's assumed bound


def _as_pairs(pm):
    return np.stack([pm.a.astype(np.int64), pm.b.astype(np.int64)], axis=1)


def _check(ctx, a, b, ratio, label=""):
    want = oracle.match_pair(a, b, ratio)
    for exact in (True, False):
        set_mode(ctx, exact)
        try:
            got = _as_pairs(match_pair(a, b, ratio, ctx=ctx))
        finally:
            set_mode(ctx, False)
        assert got.shape == want.shape, (label, exact, got.shape, want.shape)
        np.testing.assert_array_equal(got, want, err_msg="%s exact=%s" % (label, exact))
    return len(want)


@pytest.mark.parametrize("na,nb", [(1, 1), (2, 3), (63, 64), (64, 65), (129, 70), (300, 257)])
def test_match_pair_ragged(ctx, na, nb):
    fr = synthetic_frames(2, max(na, nb), seed=na * 1000 + nb, overlap=0.7)
    a, b = fr[0][:na], fr[1][:nb]
    for ratio in (0.75, 0.8, 1.0):
        _check(ctx, a, b, ratio, "%dx%d r=%g" % (na, nb, ratio))


def test_match_pair_empty(ctx):
    a = sift_like(np.random.default_rng(0), 10)
    z = np.zeros((0, 128), np.float32)
    assert len(match_pair(a, z, ctx=ctx)) == 0
    assert len(match_pair(z, a, ctx=ctx)) == 0
    assert len(match_pair(z, z, ctx=ctx)) == 0


def test_match_pair_ties_and_duplicates(ctx):
    rng = np.random.default_rng(7)
    # grid components: exact distances, many exact ties across tiles
    a = (rng.integers(0, 3, (150, 128)) / 64.0).astype(np.float32)
    b = np.concatenate([a[rng.permutation(150)[:90]], a[:40], (rng.integers(0, 3, (30, 128)) / 64.0)])
    b = np.ascontiguousarray(b[rng.permutation(len(b))], np.float32)
    for ratio in (0.75, 1.0):
        _check(ctx, a, b, ratio, "ties r=%g" % ratio)
    # identical sets: distance 0 matches, second = 0 keeps them only when
    # duplicated (0 <= r * 0)
    dup = np.concatenate([a[:70], a[:70]])
    _check(ctx, a[:70], dup, 0.75, "dup")


def test_match_pair_nan_rows(ctx):
    fr = synthetic_frames(2, 200, seed=9, overlap=0.8)
    a, b = fr[0].copy(), fr[1].copy()
    a[3, 7] = np.nan
    a[150, 0] = np.nan
    b[64, 127] = np.nan
    _check(ctx, a, b, 0.8, "nan")


def test_match_frames_batch(ctx):
    counts = [300, 0, 257, 64, 1, 190]
    frames = [f[:n] for f, n in zip(synthetic_frames(len(counts), 300, seed=3, overlap=0.6), counts)]
    fm = FrameMatcher(ctx, frames)
    pairs = job_pairs(len(frames), sequential=False) + [(2, 2), (5, 0)]
    res = fm.match(pairs, 0.8)
    tot = 0
    for (i, j), pm in zip(pairs, res):
        want = oracle.match_pair(frames[i], frames[j], 0.8)
        got = np.stack([pm.a.astype(np.int64), pm.b.astype(np.int64)], axis=1)
        np.testing.assert_array_equal(got, want, err_msg="pair (%d, %d)" % (i, j))
        tot += len(want)
    assert tot > 100
    st = fm.stats()
    assert st["pairs"] == len(pairs) and st["matches"] == tot
    # sequential pair list of run_sfm
    seq = job_pairs(len(frames), sequential=True)
    for (i, j), pm in zip(seq, fm.match(seq, 0.75)):
        want = oracle.match_pair(frames[i], frames[j], 0.75)
        np.testing.assert_array_equal(np.stack([pm.a, pm.b], 1).astype(np.int64), want)


def test_match_pair_bad_args(ctx):
    import ctypes as C
    from rustcolmap_amd._lib import MvsError, check
    fm = FrameMatcher(ctx, synthetic_frames(2, 10, seed=1))
    with pytest.raises(MvsError):
        fm.match([(0, 2)])
    with pytest.raises(MvsError):
        fm.match([(-1, 0)])
    with pytest.raises(MvsError):
        check(ctx._lib.mvs_match_set_mode(ctx._h, 2), "mvs_match_set_mode")
    counts = np.array([3, -1], np.int32)
    d = np.zeros((3, 128), np.float32)
    with pytest.raises(MvsError):
        check(ctx._lib.mvs_upload_descriptors(ctx._h, d.ctypes.data_as(C.c_void_p),
                                              counts.ctypes.data_as(C.c_void_p), 2), "mvs_upload_descriptors")
    tiny = synthetic_frames(2, 8, seed=2)
    tiny[0] = tiny[0].copy()
    tiny[0][0, 0] = 1e-30   # outside [2^-40, 2^40]: the bound check refuses the set
    with pytest.raises(MvsError):
        bound_check(ctx, tiny[0], tiny[1])
    # the context still works after the errors
    fm2 = FrameMatcher(ctx, synthetic_frames(2, 100, seed=3))
    res = fm2.match([(0, 1)], 0.75)
    fr = synthetic_frames(2, 100, seed=3)
    np.testing.assert_array_equal(_as_pairs(res[0]), oracle.match_pair(fr[0], fr[1], 0.75))


def test_match_fullsize_2048(ctx):
    """run_sfm's max_features = 2048 per frame: the full oracle at one pair,
    symmetry (match(b, a) = match(a, b) swapped) at the batch."""
    frames = synthetic_frames(4, 2048, seed=21, overlap=0.5)
    want = oracle.match_pair(frames[0], frames[1], 0.75)
    n = _check(ctx, frames[0], frames[1], 0.75, "2048")
    assert n > 500
    fm = FrameMatcher(ctx, frames)
    pairs = job_pairs(4, sequential=False)
    fwd = fm.match(pairs, 0.75)
    bwd = fm.match([(j, i) for i, j in pairs], 0.75)
    np.testing.assert_array_equal(np.stack([fwd[0].a, fwd[0].b], 1).astype(np.int64), want)
    for f, r in zip(fwd, bwd):
        sw = np.stack([r.b, r.a], 1)
        sw = sw[np.argsort(sw[:, 0], kind="stable")]
        np.testing.assert_array_equal(np.stack([f.a, f.b], 1), sw)


def test_filter_bound_margin(ctx):
    """The measured |exact - estimate| / (|a|^2 + |b|^2) stays far below the
    bound the filter assumes, on SIFT-shaped sets and on wide-range ones."""
    rng = np.random.default_rng(2)
    fr = synthetic_frames(2, 600, seed=5, overlap=0.5)
    worst = bound_check(ctx, fr[0], fr[1])
    assert 0.0 < worst < EPS_REL / 8, worst
    # full-mantissa components over 24 binades, signs mixed
    a = (rng.standard_normal((300, 128)) * 2.0 ** rng.integers(-12, 12, (300, 1))).astype(np.float32)
    b = np.concatenate([a[:150] * np.float32(1.0 + 2.0 ** -10), (rng.standard_normal((170, 128)) * 64).astype(np.float32)])
    worst2 = bound_check(ctx, a, b)
    assert worst2 < EPS_REL / 8, worst2


def test_unsafe_sets_take_exact_tiles(ctx):
    fr = synthetic_frames(3, 300, seed=8, overlap=0.6)
    fr[1] = fr[1].copy()
    fr[1][7, 3] = 1e-20     # below 2^-40: frame 1 cannot use the filter
    fm = FrameMatcher(ctx, fr)
    pairs = [(0, 1), (1, 2), (0, 2)]
    res = fm.match(pairs, 0.8)
    st = fm.stats()
    assert st["filtered_pairs"] == 1
    for (i, j), pm in zip(pairs, res):
        np.testing.assert_array_equal(_as_pairs(pm), oracle.match_pair(fr[i], fr[j], 0.8))


def test_filter_overflow_rows(ctx):
    """Clusters of 20 near-identical descriptors give rows with more than 16
    candidates: those are scanned in full."""
    rng = np.random.default_rng(4)
    base = sift_like(rng, 24)
    a = np.repeat(base, 20, axis=0) + (rng.standard_normal((480, 128)) * 1e-4).astype(np.float32)
    a = np.abs(a).astype(np.float32)
    b = np.concatenate([a[rng.permutation(480)[:200]], sift_like(rng, 100)])
    b = np.ascontiguousarray(b + (rng.standard_normal(b.shape) * 1e-5).astype(np.float32))
    b = np.abs(b).astype(np.float32)
    _check(ctx, a, b, 0.8, "clusters")
    match_pair(a, b, 0.8, ctx=ctx)
    assert match_stats(ctx)["overflow_items"] > 0


def test_match_large_sets_and_long_sequence(ctx):
    """Sets past 2048 (6000 and 5000 descriptors: 47 x 40 filter tiles,
    ragged) against the oracle; then a 60-frame sequential run (174 pairs)
    checked on sampled pairs and by symmetry on all of them."""
    big = synthetic_frames(2, 6000, seed=31, overlap=0.4)
    n = _check(ctx, big[0], big[1][:5000], 0.75, "6000x5000")
    assert n > 1000
    frames = synthetic_frames(60, 2048, seed=32, overlap=0.6)
    fm = FrameMatcher(ctx, frames)
    pairs = job_pairs(60, sequential=True)
    fwd = fm.match(pairs, 0.75)
    assert fm.stats()["filtered_pairs"] == len(pairs)
    for p in (0, 85, len(pairs) - 1):
        i, j = pairs[p]
        np.testing.assert_array_equal(_as_pairs(fwd[p]), oracle.match_pair(frames[i], frames[j], 0.75))
    bwd = fm.match([(j, i) for i, j in pairs], 0.75)
    for f, r in zip(fwd, bwd):
        sw = np.stack([r.b, r.a], 1)
        sw = sw[np.argsort(sw[:, 0], kind="stable")]
        np.testing.assert_array_equal(np.stack([f.a, f.b], 1), sw)
