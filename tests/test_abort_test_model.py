"""The device's row test for plane_cost's early abort, against the
reference's own arithmetic.

stereo.rs:215-220 aborts when the f32 quotient RN(sum / D) exceeds
thr = hint + 0.12 (D = rows_done * 5 * (si + 1), exact in f32).  The device
kernels decide it without a divide: with m the f64 midpoint of thr and its
f32 successor, RN(sum / D) > thr  <=>  sum > m * D, one f64 compare (the tie
sum == m * D cannot occur for a normal thr, since m * D then has an odd part
of at least 25 bits).  The check runs numpy's correctly rounded f32 divide
on sums placed at and around the boundary, where a wrong midpoint, a
missing tie rule or a rounding slip would flip the answer.
"""
import numpy as np

F32 = np.float32


def _midpoint(thr):
    t = thr.astype(np.float32)
    succ = np.nextafter(t, F32(np.inf))
    return 0.5 * (t.astype(np.float64) + succ.astype(np.float64))


def _thresholds(rng, n):
    # thr = (best + 0.02) + 0.12 in f32, best a cost in [0, 1]
    best = np.concatenate([rng.random(n).astype(F32), np.array([0.0, 1.0, 0.5, 0.25], F32)])
    return (best + F32(0.02)) + F32(0.12)


def _divisors(max_k):
    d = set()
    for si in range(max_k):
        for r in range(5):
            d.add(5 * (r + 1) * (si + 1))
    return np.array(sorted(d), np.float64)


def _sums_around(x, k):
    """f32 values within k units in the last place of x (f64)."""
    s0 = x.astype(F32)
    out = [s0]
    up, dn = s0, s0
    for _ in range(k):
        up = np.nextafter(up, F32(np.inf))
        dn = np.nextafter(dn, F32(0))
        out += [up, dn]
    return np.stack(out)


def test_row_test_equals_the_f32_quotient_test_at_the_boundary():
    rng = np.random.default_rng(7)
    thr = _thresholds(rng, 4000)
    m = _midpoint(thr)
    for D in _divisors(12):
        x = m * D                                  # exact in f64
        s = _sums_around(x, 3)                     # (7, n) f32 sums at the boundary
        ref = (s / F32(D)).astype(F32) > thr       # stereo.rs:217, f32 divide
        dev = s.astype(np.float64) > x
        assert np.array_equal(ref, dev), D
        # the tie the rule would have to break never occurs
        assert not np.any(s.astype(np.float64) == x), D
        # both answers occur at every divisor, so the boundary is really probed
        assert ref.any() and not ref.all()


def test_row_test_on_random_sums_and_large_source_counts():
    rng = np.random.default_rng(11)
    thr = _thresholds(rng, 2000)
    m = _midpoint(thr)
    for D in np.concatenate([_divisors(4), np.array([25.0 * 64, 25.0 * 4096, 25.0 * 600000])]):
        s = (rng.random((8, thr.size)) * (2.0 * float(D) * thr)).astype(F32)
        s = np.concatenate([s, _sums_around(m * D, 2)])
        ref = (s / F32(D)).astype(F32) > thr
        dev = s.astype(np.float64) > m * D
        assert np.array_equal(ref, dev), D


def test_infinite_threshold_never_aborts():
    # the init sweep's first hypothesis: best = inf, thr = inf; the device
    # forms the successor by incrementing the bit pattern, 0x7f800001 = NaN
    thr = np.array([np.inf], F32)
    succ = (thr.view(np.uint32) + np.uint32(1)).view(F32)
    assert np.isnan(succ[0])
    with np.errstate(invalid="ignore"):
        m = 0.5 * (thr.astype(np.float64) + succ.astype(np.float64))
    assert np.isnan(m[0])
    s = np.array([0.0, 1.0, 200.0, 3.4e38], np.float64)
    assert not np.any(s > m[0] * 5.0)
