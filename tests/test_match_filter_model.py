"""CPU model of the filtered descriptor search (no GPU): the bf16 hi/lo
split, the product estimate, the per-row threshold T = second smallest upper
bound and the candidate sets L <= T.  Checks on SIFT-shaped and wide-range
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


def stats_filtered(S, t, A):
    U = (np.float32(1 + EPS_REL) * t - np.float32(2) * ((t - A) / np.float32(2))).astype(np.float32)
    L = (np.float32(1 - EPS_REL) * t - np.float32(2) * ((t - A) / np.float32(2))).astype(np.float32)
    T = np.partition(U, 1, axis=1)[:, 1] if U.shape[1] > 1 else np.full(len(U), np.inf, np.float32)
    out = []
    ncand = []
    for i in range(len(S)):
        c = np.nonzero(L[i] <= T[i])[0]
        ncand.append(len(c))
        v = S[i, c]
        k = np.lexsort((c, v))[0]
        b1, bi = v[k], c[k]
        b2 = np.partition(v, 1)[1] if len(c) > 1 else np.inf
        out.append((bi, b1, b2))
    bi, b1, b2 = map(np.array, zip(*out))
    return bi, b1, b2, np.array(ncand)


def _sets():
    rng = np.random.default_rng(0)
    fr = synthetic_frames(2, 300, seed=1, overlap=0.5)
    yield "sift", fr[0], fr[1]
    a = (rng.standard_normal((200, 128)) * 2.0 ** rng.integers(-12, 12, (200, 1))).astype(np.float32)
    b = np.concatenate([a[:100] * np.float32(1 + 2.0 ** -10), (rng.standard_normal((120, 128)) * 8).astype(np.float32)])
    yield "wide", a, np.ascontiguousarray(b, np.float32)
    base = sift_like(rng, 20)
    c = np.abs(np.repeat(base, 10, axis=0) + (rng.standard_normal((200, 128)) * 1e-4)).astype(np.float32)
    yield "clusters", c, np.ascontiguousarray(c[rng.permutation(200)] + np.float32(1e-5), np.float32)


@pytest.mark.parametrize("name,a,b", list(_sets()), ids=lambda v: v if isinstance(v, str) else "")
def test_filter_model(name, a, b):
    S = seq_f32(a, b)
    t, A = estimate(a, b)
    rel = np.abs(S.astype(np.float64) - A.astype(np.float64)) / t.astype(np.float64)
    assert rel.max() < EPS_REL / 8, rel.max()
    for M, tt, AA in ((S, t, A), (S.T.copy(), t.T.copy(), A.T.copy())):
        fi, f1, f2 = stats_full(M)
        gi, g1, g2, nc = stats_filtered(M, tt, AA)
        np.testing.assert_array_equal(fi, gi)
        np.testing.assert_array_equal(f1, g1)
        np.testing.assert_array_equal(f2, g2)
        if name == "sift":
            assert np.median(nc) <= 4
