"""Pins of the oracle's RNG and pass (mvs/src/stereo.rs:47-78, 271-371) by
facts that do not come from re-typing the oracle's formulas.

RNG
  * pixel_rng(x, y, salt) is one output of the textbook SplitMix64
    generator (Steele, Lea, Flood 2014; Vigna's splitmix64.c) from state
    (x*G ^ y*P2 ^ salt*P5) - G, G the golden-ratio increment; from state 0
    SplitMix64's published first outputs are 0xe220a8397b1dcdaf,
    0x6e789e6aa1b965f4, 0x06c45d188009454f, which pixel_rng(1..3, 0, 0)
    must reproduce;
  * next_u64 is Marsaglia's xorshift64 with shifts (13, 7, 17) ("Xorshift
    RNGs", J. Stat. Softw. 8(14), 2003): from 88172645463325252 the
    published sequence starts 8748534153485358512, 3040900993826735515,
    3453997556048239312.

Pass, on constructed scenes whose answer the geometry fixes (an exact
integer-disparity source, planes given instead of the init sweep):
  * T1 (double buffering): a zero-cost plane held by one pixel spreads by
    exactly one Manhattan ring per pass (stereo.rs:278, 287, 300 read the
    previous buffer);
  * T6 (draw order): a perturbation whose inv_new / zcand test fails
    consumes exactly its three g0 draws (stereo.rs:326-336), so the number
    of evaluated perturbations of a pass follows from the pinned RNG alone;
  * T7 (perturbation centre): perturbations start from the previous-pass
    plane cur even after a neighbour was accepted (stereo.rs:325, 334-336):
    on a checkerboard of a near-band plane C and a far plane N only the C
    pixels can reach the zero-cost band;
  * T8 (stored depth): the stored z is bd / a_here (stereo.rs:356-358),
    which differs from the candidate's zcand by an ulp on some pixels.
Each pin is also run against the oracle's planted mistake of that kind
(orc_run_passes variants 1-4), which it must reject.
"""
import os
import sys
from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from conftest import bits

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_golden as mg  # noqa: E402

F = np.float32
M64 = (1 << 64) - 1

# ---- RNG ------------------------------------------------------------------

GOLDEN = 0x9E3779B97F4A7C15
SPLITMIX64_FROM_0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
XORSHIFT64_SEED = 88172645463325252
XORSHIFT64_FIRST = [8748534153485358512, 3040900993826735515, 3453997556048239312]


def splitmix64(state: int):
    """Textbook SplitMix64 (Vigna's splitmix64.c): returns (output, state)."""
    state = (state + GOLDEN) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31), state


def test_splitmix64_published_vectors():
    s, out = 0, []
    for _ in range(3):
        o, s = splitmix64(s)
        out.append(o)
    assert out == SPLITMIX64_FROM_0


@pytest.mark.parametrize("impl", ["oracle", "golden"])
def test_pixel_rng_is_splitmix64(impl):
    f = oracle.pixel_rng if impl == "oracle" else mg.pixel_rng
    assert [f(k, 0, 0) for k in (1, 2, 3)] == SPLITMIX64_FROM_0
    rng = np.random.default_rng(3)
    pts = [(0, 0, 7), (-1, 5, 9), (1919, 1079, 22), (-2147483648, 2147483647, 2**64 - 1)]
    pts += [(int(a), int(b), int(c)) for a, b, c in zip(rng.integers(-5000, 5000, 50), rng.integers(-5000, 5000, 50),
                                                       rng.integers(0, 2**63, 50))]
    for x, y, salt in pts:
        st = (((x & M64) * GOLDEN) ^ ((y & M64) * 0xC2B2AE3D27D4EB4F) ^ (salt * 0x165667B19E3779F9)) & M64
        assert f(x, y, salt) == splitmix64((st - GOLDEN) & M64)[0], (x, y, salt)


def test_next_u64_is_marsaglia_xorshift64():
    import ctypes as C
    s = C.c_uint64(XORSHIFT64_SEED)
    got = [oracle.load().orc_next_u64(C.byref(s)) for _ in range(3)]
    assert got == XORSHIFT64_FIRST
    s2, out = XORSHIFT64_SEED, []
    for _ in range(3):
        s2 = mg.xorshift(s2)
        out.append(s2)
    assert out == XORSHIFT64_FIRST
    # uniform() = top 24 bits / 2^24, exactly
    st = C.c_uint64(XORSHIFT64_SEED)
    u = [oracle.load().orc_uniform(C.byref(st)) for _ in range(3)]
    assert u == [float(v >> 40) / 2**24 for v in XORSHIFT64_FIRST]


# ---- constructed pass scenes ------------------------------------------------

W, H = 48, 24
FOC = 100.0
M = 5              # integer disparity of the zero-cost plane
FB = 25.0          # focal * baseline: disparity = FB / z
ZMIN, ZMAX = 1.0, 10.0
INTERIOR_X = (5, 37)   # window and census of ref and source free of clamping


def _view(gray, cx, cy, t):
    return SimpleNamespace(w=W, h=H, f=FOC, cx=cx, cy=cy, k=0.0, R=np.eye(3), t=np.asarray(t, np.float64),
                           gray=np.ascontiguousarray(gray, np.uint8), rgb=None)


def _scene(seed=5):
    """Reference view and one source translated along x by FB / FOC.  The
    source's principal point sits half a pixel further, so a fronto plane of
    disparity d maps window pixel x to column trunc(x + d + 0.5): every
    disparity in [M - 0.5, M + 0.5) reads column x + M, where the source
    holds the reference's own pixel -- a cost of exactly 0 -- and rows map
    to themselves."""
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 256, (H, W), dtype=np.uint8)
    src = rng.integers(0, 256, (H, W), dtype=np.uint8)
    src[:, M:] = ref[:, :W - M]
    return oracle.OracleViews([_view(ref, W / 2, H / 2, (0, 0, 0)),
                               _view(src, W / 2 + 0.5, H / 2 + 0.5, (FB / FOC, 0, 0))])


def _fronto(disparity, cost=1.0):
    return np.array([0.0, 0.0, 1.0, F(FB / disparity), cost], np.float32)


def _planes(fill):
    return np.broadcast_to(fill, (H, W, 5)).copy()


def _is(planes, plane):
    return np.all(bits(planes) == bits(plane.astype(np.float32)), axis=-1)


def test_zero_cost_band_geometry():
    """The construction: fronto planes with disparity in [M-0.5, M+0.5) cost
    exactly 0 in the interior, disparities outside do not."""
    ov = _scene()
    ws = ov.workspace(0, [1], ZMIN, ZMAX)
    for disp, zero in [(M, True), (M - 0.45, True), (M + 0.45, True), (M - 0.6, False), (M + 0.6, False)]:
        for x in range(INTERIOR_X[0], INTERIOR_X[1] + 1, 4):
            c = ws.plane_cost(x, 11, 0.0, 0.0, 1.0, F(FB / disp), 10.0)
            assert (c == 0.0) == zero, (disp, x, c)


@pytest.mark.parametrize("variant,ok", [(0, True), (1, False)])
def test_t1_one_ring_per_pass(variant, ok):
    ov = _scene()
    x0, y0 = 24, 12
    p0 = _planes(_fronto(M + 5))          # far plane Q everywhere, stored cost 1
    star = _fronto(M, cost=0.0)           # the zero-cost plane P* at one pixel
    p0[y0, x0] = star
    yy, xx = np.mgrid[0:H, 0:W]
    ring = np.abs(xx - x0) + np.abs(yy - y0)
    planes, results = p0, []
    for p in range(4):
        planes, _ = ov.run_passes(0, [1], ZMIN, ZMAX, planes, p, p + 1, 8, nthreads=1, variant=variant)
        results.append(np.array_equal(_is(planes, star), ring <= p + 1))
    assert all(results) == ok, results


def _gauss_stream(x, y, salt):
    st = [mg.pixel_rng(x, y, salt)]

    def uniform():
        st[0] = mg.xorshift(st[0])
        return F((st[0] >> 40) / float(1 << 24))

    def gauss():
        u1 = uniform()
        u2 = uniform()
        u3 = uniform()
        return F(F(F(u1 + u2) + u3) - F(1.5))
    return gauss


N_PASSES = 8
PASSES = [0, 1, 2, 3]


def _sched(p):
    """(sig_inv, sig_n) of pass p of N_PASSES (stereo.rs:275-277); pass 0
    is (0.02, 0.3) exactly, later passes take libm's powf values."""
    return tuple(F(v) for v in oracle.schedule(p, N_PASSES))


def _reach(p):
    """Largest inverse-depth step of a pass-p perturbation: |g| <= 1.5."""
    return 1.5 * float(_sched(p)[0]) * (1.0 / ZMIN - 1.0 / ZMAX)


def _inv_test(z_cur, g0, sig_inv):
    """inv_new / zcand acceptance of stereo.rs:325-333 in f32."""
    inv_min, inv_max = F(F(1.0) / F(ZMAX)), F(F(1.0) / F(ZMIN))
    inv = F(F(1.0) / max(F(z_cur), F(1e-6)))
    inv_new = F(inv + F(F(F(sig_inv) * F(inv_max - inv_min)) * g0))
    if inv_new <= F(0.0):
        return False, None
    zc = F(F(1.0) / inv_new)
    return bool(F(ZMIN) * F(0.5) <= zc <= F(ZMAX) * F(2.0)), zc


@pytest.mark.parametrize("variant,ok", [(0, True), (2, False)])
def test_t6_rejected_perturbation_skips_nine_draws(variant, ok):
    ov = _scene()
    neighbours = 2 * (W - 1) * H + 2 * W * (H - 1)   # all transfers pass their tests (fronto, z in range)
    agree = []
    for p in PASSES:
        sig_inv = _sched(p)[0]
        # inv_new < 0.05 (zcand > 20) rejects a perturbation with g0 < -0.3: about 28 %
        z0 = F(1.0 / (0.05 + 0.3 * _reach(p) / 1.5))
        planes, ne = ov.run_passes(0, [1], ZMIN, ZMAX, _planes(_fronto(FB / z0)), p, p + 1, N_PASSES,
                                   variant=variant)
        predicted, rejected = 0, 0
        for y in range(H):
            for x in range(W):
                gauss = _gauss_stream(x, y, p + 7)
                for _ in range(3):
                    passed, _ = _inv_test(z0, gauss(), sig_inv)
                    if not passed:
                        rejected += 1
                        continue
                    gauss(), gauss(), gauss()                 # the normal's three gaussians
                    predicted += 1                            # |g| <= 1.5: the normal tests always pass
        assert 0.15 < rejected / (3 * W * H) < 0.45
        agree.append(int(ne[0]) == neighbours + predicted)
    assert all(agree) == ok, agree


def _checkerboard(p):
    """C: a fronto plane a quarter of a pass-p perturbation's reach below the
    zero-cost band; N: a fronto plane the band is out of reach from."""
    reach_px = _reach(p) * FB
    near = _fronto(M - 0.5 - 0.25 * reach_px)
    far = _fronto(M + 4.5)
    yy, xx = np.mgrid[0:H, 0:W]
    is_c = (xx + yy) % 2 == 0
    return np.where(is_c[..., None], near, far).astype(np.float32), is_c, near


def _interior():
    yy, xx = np.mgrid[0:H, 0:W]
    return (xx >= INTERIOR_X[0]) & (xx <= INTERIOR_X[1])


@pytest.mark.parametrize("variant,ok", [(0, True), (3, False)])
def test_t7_perturbations_centre_on_cur(variant, ok):
    ov = _scene()
    res = []
    for p in PASSES:
        p0, is_c, _ = _checkerboard(p)
        planes, _ = ov.run_passes(0, [1], ZMIN, ZMAX, p0, p, p + 1, N_PASSES, variant=variant)
        zero = (planes[..., 4] == 0.0) & _interior()
        frac_c = zero[is_c & _interior()].mean()
        frac_n = zero[~is_c & _interior()].mean()
        # every C pixel first accepts a neighbour's N plane (cost < 1), so only
        # perturbations around cur = C can reach the band
        res.append((frac_c > 0.15 and frac_n == 0.0, frac_c, frac_n))
    assert all(r[0] for r in res) == ok, res


@pytest.mark.parametrize("variant,ok", [(0, True), (4, False)])
def test_t8_stored_depth_is_bd_over_a(variant, ok):
    ov = _scene()
    rx, ry, _ = ov.workspace(0, [1], ZMIN, ZMAX).export()
    checked = differs = matches = 0
    for p in PASSES:
        sig_inv, sig_n = _sched(p)
        p0, is_c, near = _checkerboard(p)
        planes, _ = ov.run_passes(0, [1], ZMIN, ZMAX, p0, p, p + 1, N_PASSES, variant=variant)
        for y, x in np.argwhere((planes[..., 4] == 0.0) & is_c & _interior()):
            gauss = _gauss_stream(int(x), int(y), p + 7)
            got = planes[y, x]
            for _ in range(3):
                passed, zc = _inv_test(near[3], gauss(), sig_inv)
                if not passed:
                    continue
                n = [F(F(near[j]) + F(sig_n * gauss())) for j in range(3)]
                ln = F(np.sqrt(F(F(F(n[0] * n[0]) + F(n[1] * n[1])) + F(n[2] * n[2]))))
                n = [F(v / ln) for v in n]
                if not all(bits(np.array(n, np.float32)) == bits(got[:3])):
                    continue
                # the accepted perturbation (its normal is stored unchanged)
                a = F(F(F(n[0] * rx[y, x]) + F(n[1] * ry[y, x])) + n[2])
                bd = F(zc * a)
                z_expected = F(bd / a)
                checked += 1
                differs += int(z_expected != zc)
                matches += int(bits(np.array([z_expected]))[0] == bits(got[3:4])[0])
                break
    assert checked > 40 and differs > 0, (checked, differs)
    assert (matches == checked) == ok, (matches, checked, differs)
