"""CPU model of the filtered descriptor search (no GPU): the bf16 hi/lo
split, the product estimate, the per-tile records (two smallest upper bounds,
a lower bound of the tile's L, the mask of elements with L <= the tile's
second smallest U), the row threshold T = second smallest upper bound and the
candidate sets.  Checks on SIFT-shaped and wide-range
sets that (1) the estimate stays within the assumed bound 2^-11 (|a|^2+|b|^2)
of the reference's f32 sum with a wide margin, and (2) the exact nearest /
second-nearest over the candidates equal the full scan's, for every row and
column -- the argument the device search rests on."""
import numpy as np
import pytest

from rustcolmap_amd.matching import sift_like, synthetic_frames

EPS_REL = 2.0 ** -11


def bf16_rn(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def split(x):
    hi = bf16_rn(x)
    lo = bf16_rn((x - hi).astype(np.float32))
    return hi, lo


def seq_f32(a, b):
    s = np.zeros((len(a), len(b)), np.float32)
    for k in range(128):
        d = (a[:, None, k] - b[None, :, k]).astype(np.float32)
        s = (s + (d * d).astype(np.float32)).astype(np.float32)
    return s


def estimate(a, b):
    """f32 result of the three bf16 products (exact products, f64 sums
    rounded once: one admissible accumulation)."""
    ah, al = split(a)
    bh, bl = split(b)
    g = (ah.astype(np.float64) @ bh.T.astype(np.float64) + ah.astype(np.float64) @ bl.T.astype(np.float64)
         + al.astype(np.float64) @ bh.T.astype(np.float64)).astype(np.float32)
    na = (a.astype(np.float32) ** 2).sum(1, dtype=np.float32)
    nb = (b.astype(np.float32) ** 2).sum(1, dtype=np.float32)
    t = (na[:, None] + nb[None, :]).astype(np.float32)
    return t, (t - np.float32(2) * g).astype(np.float32)


def stats_full(S):
    bi = np.argmin(S, axis=1)
    b1 = S[np.arange(len(S)), bi]
    b2 = np.partition(S, 1, axis=1)[:, 1] if S.shape[1] > 1 else np.full(len(S), np.inf, np.float32)
    return bi, b1, b2


def stats_filtered(S, t, A, tile=128, cap=16):
    """The device search, row by row: per tile of `tile` columns the two
    smallest upper bounds U, a lower bound lb of the tile's L values and the
    mask of elements with L <= the tile's u2; T = the second smallest U of the
    row; candidates = the masked elements of the tiles with lb <= T (more
    than `cap`: the row is scanned in full)."""
    eps = np.float32(EPS_REL)
    tU = (np.float32(1) + eps) * t
    U = (tU - np.float32(2) * ((t - A) / np.float32(2))).astype(np.float32)
    kb = np.float32(2) * eps / (np.float32(1) + eps)
    L = (U - kb * tU).astype(np.float32)
    out = []
    ncand = []
    nfull = 0
    n = S.shape[1]
    for i in range(len(S)):
        tiles = []
        for c0 in range(0, n, tile):
            u = np.sort(U[i, c0:c0 + tile])
            u1, u2 = u[0], (u[1] if len(u) > 1 else np.float32(np.inf))
            lb = u1 - kb * tU[i, c0:c0 + tile].max() * np.float32(1 + 2.0 ** -20)
            assert lb <= L[i, c0:c0 + tile].min()
            tiles.append((u1, u2, lb, c0 + np.nonzero(L[i, c0:c0 + tile] <= u2)[0]))
        T = np.sort(np.array([x for tt in tiles for x in tt[:2]], np.float32))[1]
        c = np.concatenate([tt[3] for tt in tiles if tt[2] <= T]).astype(np.int64)
        assert set(np.nonzero(L[i] <= T)[0]) <= set(c)
        if len(c) > cap:
            nfull += 1
            c = np.arange(n)
        ncand.append(len(c))
        v = S[i, c]
        k = np.lexsort((c, v))[0]
        b1, bi = v[k], c[k]
        b2 = np.partition(v, 1)[1] if len(c) > 1 else np.inf
        out.append((bi, b1, b2))
    bi, b1, b2 = map(np.array, zip(*out))
    return bi, b1, b2, np.array(ncand), nfull


def _sets():
    rng = np.random.default_rng(0)
    fr = synthetic_frames(2, 300, seed=1, overlap=0.5)
    yield "sift", fr[0], fr[1]
    a = (rng.standard_normal((200, 128)) * 2.0 ** rng.integers(-12, 12, (200, 1))).astype(np.float32)
    b = np.concatenate([a[:100] * np.float32(1 + 2.0 ** -10), (rng.standard_normal((120, 128)) * 8).astype(np.float32)])
    yield "wide", a, np.ascontiguousarray(b, np.float32)
    base = sift_like(rng, 10)
    c = np.abs(np.repeat(base, 20, axis=0) + (rng.standard_normal((200, 128)) * 1e-4)).astype(np.float32)
    yield "clusters", c, np.ascontiguousarray(c[rng.permutation(200)] + np.float32(1e-5), np.float32)


@pytest.mark.parametrize("name,a,b", list(_sets()), ids=lambda v: v if isinstance(v, str) else "")
def test_filter_model(name, a, b):
    S = seq_f32(a, b)
    t, A = estimate(a, b)
    rel = np.abs(S.astype(np.float64) - A.astype(np.float64)) / t.astype(np.float64)
    assert rel.max() < EPS_REL / 8, rel.max()
    for M, tt, AA in ((S, t, A), (S.T.copy(), t.T.copy(), A.T.copy())):
        fi, f1, f2 = stats_full(M)
        gi, g1, g2, nc, nfull = stats_filtered(M, tt, AA)
        np.testing.assert_array_equal(fi, gi)
        np.testing.assert_array_equal(f1, g1)
        np.testing.assert_array_equal(f2, g2)
        if name == "sift":
            assert np.median(nc) <= 4 and nfull <= len(M) // 50
        if name == "clusters":
            assert nfull > 0   # the full-scan fallback is exercised


def _tie_components(rng, n, scale_exp):
    """Components at bf16 rounding ties: x = s (1 + 2^-8) rounds to hi = s
    (lo = 2^-8 s, the largest lo), x = s (1 + 2^-8 + 2^-17) leaves
    r = 2^-17 s (the largest residual); mixed signs and binary scales."""
    base = np.array([1.0 + 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -17, 1.0 + 3 * 2.0 ** -9, 1.0 - 2.0 ** -9], np.float64)
    x = rng.choice(base, size=(n, 128)) * rng.choice([-1.0, 1.0], size=(n, 128))
    x *= 2.0 ** rng.integers(-scale_exp, scale_exp + 1, size=(n, 1))
    return x.astype(np.float32)


def test_split_residuals_within_their_bounds():
    """bf16 keeps 8 significant bits: |lo| <= 2^-8 |x| and |r| <= 2^-17 |x|,
    both reached by the tie components."""
    rng = np.random.default_rng(9)
    x = np.concatenate([_tie_components(rng, 64, 6).ravel(), rng.normal(size=100000).astype(np.float32)])
    hi, lo = split(x)
    r = x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64)
    ax = np.abs(x.astype(np.float64))
    assert np.all(np.abs(lo) <= 2.0 ** -8 * ax)
    assert np.all(np.abs(r) <= 2.0 ** -17 * ax)
    assert np.max(np.abs(lo) / ax) > 2.0 ** -8 * 0.99        # the bounds are reached
    assert np.max(np.abs(r) / ax) > 2.0 ** -17 * 0.99


@pytest.mark.parametrize("scale_exp", [0, 5])
def test_estimate_error_on_bf16_ties(scale_exp):
    """Adversarial sets whose components sit at bf16 rounding ties (largest
    lo and residual): the estimate's error against the reference's f32 sum
    stays below 8e-5 (|a|^2 + |b|^2), a sixth of EPS_REL; identical and
    sign-flipped descriptors are included (exact distance 0 and the largest
    cross terms)."""
    rng = np.random.default_rng(10 + scale_exp)
    a = _tie_components(rng, 96, scale_exp)
    b = np.concatenate([a[:32], -a[32:48], _tie_components(rng, 48, scale_exp)])
    ref = seq_f32(a, b).astype(np.float64)
    t, A = estimate(a, b)
    rel = np.abs(A.astype(np.float64) - ref) / t.astype(np.float64)
    assert rel.max() < 8e-5 < EPS_REL / 6
