"""The allocation-free census entry point (mvs_census_noalloc) against
mvs_census, which allocates its device staging per call: equal codes on
calls in a row over images of falling and rising size (so stale scratch
contents would show), the largest accepted image, a second context on the
same GPU, and MVS_ERR_LIMIT above MVS_NOALLOC_MAX_PIXELS."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_PIXELS = 1 << 22      # MVS_NOALLOC_MAX_PIXELS


def _say(msg):
    print(msg, flush=True)


def _noalloc(ctx, g):
    from rustcolmap_amd import _lib
    h, w = g.shape
    out = np.full((h, w), 0xFFFFFFFFFFFFFFFF, np.uint64)
    _lib.check(_lib.load().mvs_census_noalloc(ctx._h, g.ctypes.data_as(C.c_void_p), int(w), int(h),
                                               out.ctypes.data_as(C.c_void_p)), "mvs_census_noalloc")
    return out


def _image(rng, h, w):
    g = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    g[: h // 3, : w // 3] = 40     # flat region: all-zero codes there
    return g


def test_census_noalloc_matches_census(ctx):
    import oracle
    rng = np.random.default_rng(8)
    for (h, w) in [(1440, 2560), (97, 131), (2048, 2048), (1, 1), (7, 3)]:
        _say("census_noalloc %dx%d" % (w, h))
        g = _image(rng, h, w)
        want = ctx.census(g)
        got = _noalloc(ctx, g)
        assert np.array_equal(got, want), (w, h)
        if h * w < 20000:
            assert np.array_equal(got, oracle.census(g)), (w, h)


def test_census_noalloc_second_context_and_limit(ctx):
    from rustcolmap_amd import _lib
    from rustcolmap_amd.mvs import MvsContext
    rng = np.random.default_rng(9)
    other = MvsContext(0)
    try:
        for (h, w) in [(480, 640), (33, 65)]:
            _say("census_noalloc second context %dx%d" % (w, h))
            g = _image(rng, h, w)
            assert np.array_equal(_noalloc(other, g), ctx.census(g)), (w, h)
    finally:
        other.close()
    # above the maximum: rejected on the host before any device work
    dummy = np.zeros(1, np.uint8)
    out = np.zeros(1, np.uint64)
    r = _lib.load().mvs_census_noalloc(ctx._h, dummy.ctypes.data_as(C.c_void_p), 2049, 2048,
                                       out.ctypes.data_as(C.c_void_p))
    assert r == -5 and 2049 * 2048 > MAX_PIXELS
    assert b"MVS_NOALLOC_MAX_PIXELS" in _lib.load().mvs_last_error()
    # the context still works afterwards
    g = _image(rng, 64, 64)
    assert np.array_equal(_noalloc(ctx, g), ctx.census(g))
