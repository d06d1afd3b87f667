"""Input stage of load_views (mvs/src/lib.rs:185-197): the restatement of
the image crate's Triangle resize and luma in oracle/image_ops.py, pinned by
closed forms (the crate itself, image 0.24.9, is not in the reference tree
and cannot be run here: parity with it is unpinned)."""
import numpy as np

from oracle import image_ops as I

F = np.float32


def test_downscale_by_2_weights_are_1_3_3_1():
    for o, (left, w) in enumerate(I._taps(16, 8)):
        if 0 < o < 7:
            assert left == 2 * o - 1 and w.tolist() == [0.125, 0.375, 0.375, 0.125]
    left, w = I._taps(16, 8)[0]                       # clamped at the border: taps 0, 1, 2
    assert left == 0 and np.allclose(w, np.array([0.75, 0.75, 0.25]) / 1.75, rtol=0, atol=1e-7)


def test_downscale_by_3_and_upscale_by_2_weights():
    left, w = I._taps(30, 10)[4]                      # centre 13.5: taps 10..16
    assert left == 10 and len(w) == 7
    assert np.allclose(w, np.array([0, 1, 2, 3, 2, 1, 0]) / 9.0, atol=1e-7) and w[0] == 0 and w[-1] == 0
    left, w = I._taps(10, 20)[2]                      # upscale: centre 1.25 -> 0.25 * p0 + 0.75 * p1
    assert left == 0 and w.tolist() == [0.25, 0.75, 0.0]


def test_resize_closed_forms():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, size=(37, 53, 3), dtype=np.uint8)
    assert np.array_equal(I.resize_triangle(img, 53, 37), img)            # equal size: a copy
    const = np.full((40, 60), 201, np.uint8)
    assert np.all(I.resize_triangle(const, 23, 17) == 201)
    rows = np.repeat(rng.integers(0, 256, size=(40, 1), dtype=np.uint8), 60, axis=1)
    out = I.resize_triangle(rows, 25, 40)                                 # only columns resampled
    assert np.array_equal(out, rows[:, :25])
    ramp = np.tile(np.arange(0, 200, 20, dtype=np.uint8), (4, 1))          # 0, 20, ..., 180 along x
    up = I.resize_triangle(ramp, 20, 4)
    # upscale 2x: interior outputs are 0.25 / 0.75 blends of neighbours
    assert up[0, 3].item() == round(0.75 * 20 + 0.25 * 40) and up[0, 4].item() == round(0.25 * 20 + 0.75 * 40)
    pairs = np.zeros((2, 2), np.uint8)
    pairs[:, 1] = 255
    # interior windows weigh 0 and 255 equally: 127.5 rounds up; the clamped
    # border windows (taps 0-2: 0.75, 0.75, 0.25 of 0, 255, 0, over 1.75;
    # taps 13-15: 0.25, 0.75, 0.75 of 255, 0, 255) give 109.3 and 145.7
    assert I.resize_triangle(np.tile(pairs, (4, 8)), 8, 4).tolist() == [[109] + [128] * 6 + [146]] * 4


def test_luma_closed_forms():
    """image 0.24.9's integer luma: every grey is its own luma (the weights
    2126 + 7152 + 722 sum to the divisor 10000), the primaries give
    floor(255 w / 10000), adding 1 to one channel never lowers the luma,
    and every value is the floor of the weighted sum / 10000 -- checked
    against exact rational arithmetic over a sample of all 2^24 colours."""
    grey = np.repeat(np.arange(256, dtype=np.uint8)[None, :, None], 3, 2)
    assert np.array_equal(I.rgb_to_luma(grey)[0], np.arange(256))
    rgb = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [0, 0, 0], [10, 10, 10]]], np.uint8)
    assert I.rgb_to_luma(rgb)[0].tolist() == [54, 182, 18, 0, 10]
    from fractions import Fraction
    rng = np.random.default_rng(2)
    c = rng.integers(0, 256, size=(1, 4000, 3), dtype=np.uint8)
    got = I.rgb_to_luma(c)[0]
    for (r, g, b), l in zip(c[0].tolist(), got.tolist()):
        assert l == int(Fraction(2126 * r + 7152 * g + 722 * b, 10000))
    for ch in range(3):
        up = c.copy()
        up[..., ch] = np.minimum(up[..., ch].astype(int) + 1, 255).astype(np.uint8)
        assert np.all(I.rgb_to_luma(up) >= I.rgb_to_luma(c))
    gray = np.arange(12, dtype=np.uint8).reshape(3, 4)
    gg, rr = I.load_view_pixels(gray, 4, 3)
    assert np.array_equal(gg, gray) and np.array_equal(rr, np.repeat(gray[:, :, None], 3, 2))
