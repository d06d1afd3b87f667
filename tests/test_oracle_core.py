"""Pins of the oracle's integer parts: RNG, census, the popcount/48 table,
camera helpers and the texture gate.

Each check is independent of the C code it tests: known-answer fixtures
from the pure-Python big-int / brute-force implementation in
tests/golden/make_golden.py, closed forms, and exhaustive enumeration.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_rng_matches_bigint_kat():
    """pixel_rng + xorshift64 + uniform (stereo.rs:47-72) vs the big-int KAT."""
    rows = json.load(open(os.path.join(GOLD, "rng_kat.json")))
    assert len(rows) >= 40
    for r in rows:
        s = oracle.pixel_rng(r["x"], r["y"], r["salt"])
        assert s == int(r["seed"]), r
        u = oracle.uniform_stream(s, 12)
        # uniform() is exactly k / 2^24 for the top 24 bits k
        assert [Fraction(x) for x in u] == [Fraction(k, 1 << 24) for k in r["u24"]], r


def test_gauss_is_irwin_hall3_in_f32():
    rows = json.load(open(os.path.join(GOLD, "rng_kat.json")))
    for r in rows[:10]:
        s = int(r["seed"])
        g = oracle.gauss_stream(s, 4)
        u = [np.float32(k / float(1 << 24)) for k in r["u24"]]
        for i in range(4):
            exp = np.float32(np.float32(np.float32(u[3 * i] + u[3 * i + 1]) + u[3 * i + 2]) - np.float32(1.5))
            assert np.float32(g[i]) == exp


def test_census_matches_bruteforce_kat():
    cases = json.load(open(os.path.join(GOLD, "census_kat.json")))
    for c in cases:
        g = np.asarray(c["gray"], np.uint8).reshape(c["h"], c["w"])
        got = oracle.census(g).reshape(-1)
        assert [int(v) for v in got] == [int(v) for v in c["census"]], (c["w"], c["h"])


def _census_numpy(g):
    """Vectorised census by padding with edge values (clamp-to-edge)."""
    h, w = g.shape
    p = np.pad(g, 3, mode="edge")
    out = np.zeros((h, w), np.uint64)
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            if dx == 0 and dy == 0:
                continue
            nb = p[3 + dy:3 + dy + h, 3 + dx:3 + dx + w]
            out |= (nb < g).astype(np.uint64) << np.uint64((dy + 3) * 7 + (dx + 3))
    return out


@pytest.mark.parametrize("shape", [(1, 1), (2, 9), (40, 33), (120, 97)])
def test_census_matches_vectorised_and_invariants(shape):
    rng = np.random.default_rng(sum(shape))
    g = rng.integers(0, 256, size=shape, dtype=np.uint8)
    got = oracle.census(g)
    assert np.array_equal(got, _census_numpy(g))
    # bit 24 (centre) and bits 49..63 never set; popcount <= 48
    assert not np.any(got & np.uint64(1 << 24))
    assert not np.any(got >> np.uint64(49))
    # constant image -> all zero; strict '<' keeps equal neighbours out
    assert not np.any(oracle.census(np.full(shape, 93, np.uint8)))


def test_census_edge_clamp_case():
    # a single bright pixel in the corner: every clamped neighbour reads it
    g = np.zeros((5, 5), np.uint8)
    g[0, 0] = 200
    c = oracle.census(g)
    # the corner pixel sees darker neighbours everywhere except the clamped
    # taps that land on itself (dx <= 0 and dy <= 0, excluding the centre)
    exp = 0
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            if (dx, dy) == (0, 0):
                continue
            if not (dx <= 0 and dy <= 0):
                exp |= 1 << ((dy + 3) * 7 + (dx + 3))
    assert int(c[0, 0]) == exp
    assert int(c[4, 4]) == 0


def _round_f32(x: Fraction) -> Fraction:
    """Exact round-to-nearest-even of a rational to binary32 (normal range)."""
    if x == 0:
        return Fraction(0)
    s = -1 if x < 0 else 1
    a = abs(x)
    e = math.floor(math.log2(a.numerator) - math.log2(a.denominator))
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    q = a / ulp
    n = q.numerator // q.denominator
    r = q - n
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return s * n * ulp


def test_pc_over_48_fma_identity_exhaustive():
    """The HIP kernel forms pc/48 as q = pc*r, rem = fma(-q, 48, pc),
    fma(rem, r, q) with r = RN(1/48).  For all 49 popcounts this equals
    the correctly rounded quotient that `count_ones() as f32 / 48.0`
    produces (stereo.rs:213)."""
    r = _round_f32(Fraction(1, 48))
    assert float(r) == float(np.float32(1.0) / np.float32(48.0))
    for pc in range(49):
        a = Fraction(pc)
        q = _round_f32(a * r)
        rem = _round_f32(-q * 48 + a)
        got = _round_f32(rem * r + q)
        assert got == _round_f32(Fraction(pc, 48)), pc
        assert float(got) == float(np.float32(pc) / np.float32(48.0))


def test_normalize_pinhole_closed_form_and_radial_fixed_point():
    # k = 0: exactly (u - cx) / f  (geom.rs:28-41)
    for (u, v) in [(0.0, 0.0), (319.0, 239.0), (12.5, 200.25)]:
        x, y = oracle.normalize(275.0, 160.0, 120.0, 0.0, u, v)
        assert x == (u - 160.0) / 275.0 and y == (v - 120.0) / 275.0
    # k != 0: two fixed-point steps of xd = xn (1 + k r^2)
    f, cx, cy, k = 500.0, 320.0, 240.0, -0.05
    u, v = 600.0, 400.0
    x, y = oracle.normalize(f, cx, cy, k, u, v)
    xn, yn = (u - cx) / f, (v - cy) / f
    for _ in range(2):
        d = 1.0 + k * (xn * xn + yn * yn)
        xn, yn = (u - cx) / (f * d), (v - cy) / (f * d)
    assert (x, y) == (xn, yn)


def test_quat_to_rot_known_rotations():
    R = oracle.quat_to_rot([1.0, 0.0, 0.0, 0.0])
    assert np.array_equal(R, np.eye(3))
    c = math.cos(math.pi / 4)
    R = oracle.quat_to_rot([c, 0.0, 0.0, c])      # 90 deg about z
    assert np.allclose(R, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)
    R = oracle.quat_to_rot([2.0, 0.0, 0.0, 0.0])  # non-unit -> renormalised
    assert np.allclose(R, np.eye(3), atol=1e-15)
    q = np.array([0.3, -0.5, 0.7, 0.1])
    q = q / np.linalg.norm(q)
    R = oracle.quat_to_rot(q)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(R) - 1.0) < 1e-12
    # geom.rs:85 renormalises only when |n| - 1 > 1e-12: a quaternion of
    # norm < 1 is used as given (the product is then not a rotation)
    qs = 0.5 * q
    R = oracle.quat_to_rot(qs)
    w, x, y, z = qs
    assert R[0, 0] == 1.0 - 2.0 * (y * y + z * z)
    assert not np.allclose(R @ R.T, np.eye(3), atol=1e-3)


def test_schedule_is_libm_powf():
    """sig_inv = 0.02 * 0.28^(4t), sig_n = 0.30 * 0.42^(4t) in f32
    (stereo.rs:275-277); powf is libm's, checked to <= 1 ulp of the
    double-precision value."""
    for n in (4, 12, 16):
        for p in range(n):
            a, b = oracle.schedule(p, n)
            t = np.float32(np.float32(p) / np.float32(n))
            e = float(np.float32(t * np.float32(4.0)))
            ea = np.float32(np.float32(0.02) * np.float32(0.28 ** e))
            eb = np.float32(np.float32(0.30) * np.float32(0.42 ** e))
            assert abs(np.float32(a) - ea) <= np.spacing(ea) * 2
            assert abs(np.float32(b) - eb) <= np.spacing(eb) * 2
    assert oracle.schedule(0, 16) == (np.float32(0.02), np.float32(0.30))
