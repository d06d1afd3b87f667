"""Pins of the oracle's plane_cost (mvs/src/stereo.rs:158-228).

Views are built with f = 1, cx = cy = 0, k = 0 and R = I, so a pixel's ray
is its integer coordinate and, for a fronto-parallel plane at depth 2 and a
source translation (2m, 0, 0), every window tap lands exactly on source
column x + m.  The expected integer Hamming sums, invalid-tap counts,
sequential f32 sums and the early-abort peak follow by brute force from the
census images, independently of the oracle's warp code.
"""
from types import SimpleNamespace

import numpy as np

import oracle


def _view(gray, t=(0.0, 0.0, 0.0), f=1.0, cx=0.0, cy=0.0):
    h, w = gray.shape
    return SimpleNamespace(w=w, h=h, f=f, cx=cx, cy=cy, k=0.0, R=np.eye(3), t=np.asarray(t, np.float64),
                           gray=np.ascontiguousarray(gray, np.uint8), rgb=None)


def _expected(ref_cen, src_cens, px, py, col_map):
    """Brute-force taps: col_map(x) -> source column (or None = invalid)."""
    h, w = ref_cen.shape
    ham = 0
    ninv = 0
    s32 = np.float32(0.0)
    peak = np.float32(-np.inf)
    for si, sc in enumerate(src_cens):
        for r, dy in enumerate(range(-2, 3)):
            y = min(max(py + dy, 0), h - 1)
            for dx in range(-2, 3):
                x = min(max(px + dx, 0), w - 1)
                u = col_map[si](x)
                if u is None or not (0 <= u < sc.shape[1]):
                    s32 = np.float32(s32 + np.float32(1.0))
                    ninv += 1
                    continue
                pc = bin(int(ref_cen[y, x]) ^ int(sc[y, u])).count("1")
                ham += pc
                s32 = np.float32(s32 + np.float32(np.float32(pc) / np.float32(48.0)))
            # normaliser rows_done * 5 * (si + 1)  (stereo.rs:216-217)
            part = np.float32(s32 / np.float32((r + 1) * 5 * (si + 1)))
            peak = max(peak, part)
    total = np.float32(25 * len(src_cens))
    return ham, ninv, s32, peak, np.float32(s32 / total)


def _setup(shifts, w=20, h=14, seed=3):
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    views = [_view(ref)]
    for m in shifts:
        # source image: the reference texture shifted by m columns plus noise
        src = np.roll(ref, m, axis=1).copy()
        src[rng.random(src.shape) < 0.1] = 17
        views.append(_view(src, t=(2.0 * m, 0.0, 0.0)))
    ov = oracle.OracleViews(views)
    return ov, views


def test_identity_source_gives_zero_cost():
    rng = np.random.default_rng(0)
    g = rng.integers(0, 256, size=(12, 16), dtype=np.uint8)
    ov = oracle.OracleViews([_view(g), _view(g)])
    ws = ov.workspace(0, [1], 1.0, 10.0)
    for (px, py) in [(0, 0), (7, 5), (15, 11), (3, 11)]:
        c, info = ws.plane_cost(px, py, 0.0, 0.0, 1.0, 2.0, np.inf, info=True)
        assert c == 0.0 and info.ham_sum == 0 and info.n_invalid == 0


def test_integer_disparity_hamming_sums_and_f32_sum():
    shifts = [2, -3]
    ov, views = _setup(shifts)
    ws = ov.workspace(0, [1, 2], 1.0, 10.0)
    ref_cen = ov.census[0]
    src_cens = [ov.census[1], ov.census[2]]
    maps = [lambda x, m=m: x + m for m in shifts]
    for (px, py) in [(0, 0), (5, 6), (10, 7), (19, 13), (2, 13)]:
        ham, ninv, s32, peak, res = _expected(ref_cen, src_cens, px, py, maps)
        c, info = ws.plane_cost(px, py, 0.0, 0.0, 1.0, 2.0, np.inf, info=True)
        assert (info.ham_sum, info.n_invalid) == (ham, ninv), (px, py)
        assert np.float32(info.full_sum) == s32
        assert np.float32(info.peak) == peak
        assert np.float32(c) == res


def test_early_abort_fires_iff_peak_exceeds_threshold():
    ov, views = _setup([1, 4, -2], seed=9)
    ws = ov.workspace(0, [1, 2, 3], 1.0, 10.0)
    ref_cen = ov.census[0]
    maps = [lambda x, m=m: x + m for m in (1, 4, -2)]
    for (px, py) in [(4, 4), (12, 8), (0, 13)]:
        ham, ninv, s32, peak, res = _expected(ref_cen, [ov.census[i] for i in (1, 2, 3)], px, py, maps)
        for delta in (-1e-3, -1e-6, 0.0, 1e-6, 1e-3):
            hint = np.float32(np.float32(peak) - np.float32(0.12) + np.float32(delta))
            thr = np.float32(hint + np.float32(0.12))
            c = ws.plane_cost(px, py, 0.0, 0.0, 1.0, 2.0, float(hint))
            if peak > thr:
                assert c == np.float32(1e9), (px, py, delta)
            else:
                assert np.float32(c) == res, (px, py, delta)


def test_plane_behind_camera_and_grazing_give_one():
    ov, views = _setup([2])
    ws = ov.workspace(0, [1], 1.0, 10.0)
    c, info = ws.plane_cost(5, 5, 0.0, 0.0, 1.0, -2.0, np.inf, info=True)   # z < 0 -> zs < 1e-3
    assert c == 1.0 and info.n_invalid == 25
    c, info = ws.plane_cost(5, 5, 0.0, 0.0, -1.0, 2.0, np.inf, info=True)   # denom < 0.05
    assert c == 1.0 and info.n_invalid == 25
    c, info = ws.plane_cost(5, 5, 0.0, 0.0, 0.0499, 2.0, np.inf, info=True)
    assert c == 1.0 and info.n_invalid == 25


def test_fully_off_image_gives_one():
    ov, views = _setup([100])
    ws = ov.workspace(0, [1], 1.0, 10.0)
    c, info = ws.plane_cost(8, 8, 0.0, 0.0, 1.0, 2.0, np.inf, info=True)
    assert c == 1.0 and info.n_invalid == 25


def test_u_in_minus_one_to_zero_truncates_to_column_zero():
    """`(x) as i32` truncates toward zero: u in (-1, 0) lands on column 0 and
    is a valid tap (stereo.rs:205-207)."""
    rng = np.random.default_rng(4)
    ref = rng.integers(0, 256, size=(10, 12), dtype=np.uint8)
    src = rng.integers(0, 256, size=(10, 12), dtype=np.uint8)
    # translation -1 at depth 2: x_n = x - 0.5 exactly
    ov = oracle.OracleViews([_view(ref), _view(src, t=(-1.0, 0.0, 0.0))])
    ws = ov.workspace(0, [1], 1.0, 10.0)
    maps = [lambda x: 0 if x == 0 else x - 1]
    for (px, py) in [(0, 3), (1, 0), (6, 9)]:
        ham, ninv, s32, peak, res = _expected(ov.census[0], [ov.census[1]], px, py, maps)
        c, info = ws.plane_cost(px, py, 0.0, 0.0, 1.0, 2.0, np.inf, info=True)
        assert (info.ham_sum, info.n_invalid) == (ham, ninv)
        assert np.float32(c) == res
    # the floor() reading would mark the x = 0 taps invalid instead
    c, info = ws.plane_cost(0, 3, 0.0, 0.0, 1.0, 2.0, np.inf, info=True)
    assert info.n_invalid == 0


def test_workspace_warp_is_relative_pose():
    """rs = R_s R_r^T, ts = t_s - rs t_r (stereo.rs:131-132), cast to f32."""
    rng = np.random.default_rng(2)
    g = rng.integers(0, 256, size=(6, 8), dtype=np.uint8)

    def rot(a, b):
        ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
        return np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]]) @ np.array([[1, 0, 0], [0, cb, -sb], [0, sb, cb]])

    Rr, Rs = rot(0.3, -0.1), rot(-0.2, 0.25)
    tr, ts = np.array([0.5, -0.2, 3.0]), np.array([-1.0, 0.1, 2.5])
    vr = _view(g, t=tr, f=300.0, cx=4.0, cy=3.0)
    vr.R = Rr
    vs = _view(g, t=ts, f=310.0, cx=4.0, cy=3.0)
    vs.R = Rs
    ws = oracle.OracleViews([vr, vs]).workspace(0, [1], 1.0, 10.0)
    rx, ry, sp = ws.export()
    rs = Rs @ Rr.T
    assert np.allclose(sp[0, :9], rs.reshape(9).astype(np.float32), atol=1e-7)
    assert np.allclose(sp[0, 9:12], (ts - rs @ tr).astype(np.float32), atol=1e-6)
    assert sp[0, 12] == np.float32(310.0)
    # rays: f32 of (x - cx) / f
    assert rx[0, 0] == np.float32((0 - 4.0) / 300.0) and ry[5, 7] == np.float32((5 - 3.0) / 300.0)
