"""GPU input stage (mvs_resize_image) against the CPU restatement of the
image crate's Triangle resize and luma (oracle/image_ops.py), bit for bit:
RGB and gray, downscales by integer and fractional ratios, an upscale and
equal size; and views_from_reconstruction on native-size images."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ow,oh,mw,mh", [(640, 480, 320, 240), (1000, 750, 400, 300), (333, 222, 128, 85),
                                         (100, 80, 160, 128), (64, 48, 64, 48), (2560, 1440, 1600, 900)])
def test_resize_matches_restatement(ctx, ow, oh, mw, mh):
    from oracle import image_ops as I
    from rustcolmap_amd.mvs import resize_image
    rng = np.random.default_rng(ow + oh)
    rgb = rng.integers(0, 256, size=(oh, ow, 3), dtype=np.uint8)
    rgb[: oh // 2] = (rgb[: oh // 2] // 16) * 16          # smooth-ish half: many exact halves in the blends
    g, c = resize_image(ctx, rgb, mw, mh)
    eg, ec = I.load_view_pixels(rgb, mw, mh)
    assert np.array_equal(c, ec) and np.array_equal(g, eg)
    gray = rgb[:, :, 1].copy()
    g, c = resize_image(ctx, gray, mw, mh)
    eg, ec = I.load_view_pixels(gray, mw, mh)
    assert np.array_equal(g, eg) and np.array_equal(c, ec)


def test_views_from_native_images(ctx):
    from oracle import image_ops as I
    from rustcolmap_amd import formats as F
    rng = np.random.default_rng(5)
    cam = F.Camera(1, F.SIMPLE_RADIAL, 800, 600, [700.0, 400.0, 300.0, 0.0])
    imgs = [F.Image(i + 1, (1.0, 0.0, 0.0, 0.0), (0.1 * i, 0.0, 0.0), 1, "v%d.png" % i) for i in range(2)]
    rec = F.Reconstruction([cam], imgs, [])
    native = {im.name: (None, rng.integers(0, 256, size=(600, 800, 3), dtype=np.uint8)) for im in imgs}
    views, _, _ = F.views_from_reconstruction(rec, native, 400, ctx=ctx)
    for v, im in zip(views, imgs):
        eg, ec = I.load_view_pixels(native[im.name][1], 400, 300)
        assert (v.w, v.h) == (400, 300) and np.array_equal(v.gray, eg) and np.array_equal(v.rgb, ec)
        assert v.cam.f == 350.0 and v.cam.cx == 200.0 and v.cam.cy == 150.0


def test_luma_every_colour(ctx):
    """Every one of the 2^24 RGB colours (a 4096 x 4096 image, resized at
    equal size, which is a copy) gets the integer luma of the restatement."""
    from oracle import image_ops as I
    from rustcolmap_amd.mvs import resize_image
    v = np.arange(1 << 24, dtype=np.uint32)
    rgb = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], axis=-1).astype(np.uint8).reshape(4096, 4096, 3)
    g, c = resize_image(ctx, rgb, 4096, 4096)
    assert np.array_equal(c, rgb)
    assert np.array_equal(g, I.rgb_to_luma(rgb))
    grey = rgb[..., 0] == rgb[..., 1]
    grey &= rgb[..., 1] == rgb[..., 2]
    assert np.array_equal(g[grey], rgb[..., 0][grey])
