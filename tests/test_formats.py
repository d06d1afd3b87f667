"""Formats on either side of the dense path (rustcolmap_amd/formats.py and
the library's dense_web.json writer).

Pins, independent of the code under test:
* COLMAP binary files: expected bytes assembled field by field with
  struct.pack in the writer order of colmap-io/src/lib.rs:319-378;
* COLMAP text files: the expected text written out literally, headers as in
  colmap-io/src/lib.rs:423-490, floats as Rust's `{}` prints them (known
  answers: shortest round-trip digits, no exponent, "-0", integral values
  without a fraction);
* reader quirks of colmap-io/src/lib.rs:492-575 (comment lines, an image line
  with fewer than 10 tokens skips its pair, names with spaces, incomplete
  triples ignored, `as` wrapping of out-of-range integers, sort by id);
* cam_from_rec per camera model and the working size in closed form
  (mvs/src/lib.rs:103-117, 154-157); quat_to_rot on exact rotations and the
  reference's no-renormalisation quirk below norm 1;
* dense_web.json on clouds whose centroid, median and 95th percentile are
  known, and the C ABI writer byte-identical to the Python one;
* the depth visualisation's percentiles, polarity and invalid pixels in
  closed form, PNG structure and round trip.
"""
import ctypes as C
import json
import math
import os
import struct
import zlib

import numpy as np
import pytest

from rustcolmap_amd import formats as F


# --------------------------------------------------------------------------
# Rust float printing
# --------------------------------------------------------------------------

@pytest.mark.parametrize("x,s", [
    (1.0, "1"), (0.0, "0"), (-0.0, "-0"), (0.1, "0.1"), (-2.5, "-2.5"), (100.0, "100"),
    (1e-7, "0.0000001"), (1.25e-5, "0.0000125"), (1e16, "10000000000000000"),
    (1e21, "1000000000000000000000"), (123456789.125, "123456789.125"),
    (0.30000000000000004, "0.30000000000000004"), (1.7976931348623157e308, "17976931348623157" + "0" * 292),
    (5e-324, "0." + "0" * 323 + "5"), (float("inf"), "inf"), (float("-inf"), "-inf"), (float("nan"), "NaN"),
    (2.0 ** 60, "1152921504606847000"), (1.0 / 3.0, "0.3333333333333333"),
])
def test_rust_display_of_f64(x, s):
    assert F.rust_f64(x) == s


def test_rust_display_round_trips():
    rng = np.random.default_rng(0)
    for v in np.concatenate([rng.normal(size=200) * 10.0 ** rng.integers(-30, 30, 200),
                             rng.integers(-10 ** 6, 10 ** 6, 50).astype(np.float64)]):
        s = F.rust_f64(float(v))
        assert "e" not in s and float(s) == float(v)


# --------------------------------------------------------------------------
# a small reconstruction
# --------------------------------------------------------------------------

def _rec():
    cams = [F.Camera(1, F.SIMPLE_RADIAL, 640, 480, [500.0, 320.0, 240.0, 0.01]),
            F.Camera(2, F.PINHOLE, 800, 600, [600.0, 610.0, 400.5, 299.5])]
    imgs = [F.Image(1, (1.0, 0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1, "a.png",
                    np.array([[10.5, 20.25], [30.0, 40.0]]), np.array([7, -1], np.int64)),
            F.Image(2, (0.5, 0.5, 0.5, 0.5), (1.0, -2.0, 0.5), 2, "dir/b c.jpg")]
    pts = [F.Point3D(7, (0.1, -0.2, 3.0), (255, 0, 17), 0.75, np.array([[1, 0], [2, 5]], np.int64))]
    return F.Reconstruction(cams, imgs, pts)


def _expected_bin():
    cams = struct.pack("<Q", 2)
    cams += struct.pack("<iiQQ", 1, 2, 640, 480) + struct.pack("<4d", 500.0, 320.0, 240.0, 0.01)
    cams += struct.pack("<iiQQ", 2, 1, 800, 600) + struct.pack("<4d", 600.0, 610.0, 400.5, 299.5)
    imgs = struct.pack("<Q", 2)
    imgs += struct.pack("<i", 1) + struct.pack("<4d", 1, 0, 0, 0) + struct.pack("<3d", 0, 0, 0) + struct.pack("<i", 1)
    imgs += bytes([5]) + b"a.png" + struct.pack("<Q", 2)
    imgs += struct.pack("<ddq", 10.5, 20.25, 7) + struct.pack("<ddq", 30.0, 40.0, -1)
    imgs += struct.pack("<i", 2) + struct.pack("<4d", .5, .5, .5, .5) + struct.pack("<3d", 1, -2, .5)
    imgs += struct.pack("<i", 2) + bytes([11]) + b"dir/b c.jpg" + struct.pack("<Q", 0)
    p3 = struct.pack("<Q", 1) + struct.pack("<q", 7) + struct.pack("<3d", 0.1, -0.2, 3.0) + bytes([255, 0, 17])
    p3 += struct.pack("<d", 0.75) + struct.pack("<Q", 2) + struct.pack("<iiii", 1, 0, 2, 5)
    return cams, imgs, p3


def _same(a: F.Reconstruction, b: F.Reconstruction):
    assert [(c.camera_id, c.model_id, c.width, c.height, list(c.params)) for c in a.cameras] == \
           [(c.camera_id, c.model_id, c.width, c.height, list(c.params)) for c in b.cameras]
    assert len(a.images) == len(b.images)
    for x, y in zip(a.images, b.images):
        assert (x.image_id, tuple(x.qvec), tuple(x.tvec), x.camera_id, x.name) == \
               (y.image_id, tuple(y.qvec), tuple(y.tvec), y.camera_id, y.name)
        assert np.array_equal(x.points2d.reshape(-1, 2), y.points2d.reshape(-1, 2))
        assert np.array_equal(x.point3d_ids, y.point3d_ids)
    assert len(a.points3d) == len(b.points3d)
    for x, y in zip(a.points3d, b.points3d):
        assert (x.point3d_id, tuple(x.xyz), tuple(x.rgb), x.error) == (y.point3d_id, tuple(y.xyz), tuple(y.rgb),
                                                                       y.error)
        assert np.array_equal(np.asarray(x.track).reshape(-1, 2), np.asarray(y.track).reshape(-1, 2))


def test_binary_bytes_and_round_trip(tmp_path):
    F.write_bin_dir(_rec(), str(tmp_path))
    ec, ei, ep = _expected_bin()
    assert (tmp_path / "cameras.bin").read_bytes() == ec
    assert (tmp_path / "images.bin").read_bytes() == ei
    assert (tmp_path / "points3D.bin").read_bytes() == ep
    _same(F.read_bin_dir(str(tmp_path)), _rec())
    assert F.read_dir_any(str(tmp_path)).images[1].name == "dir/b c.jpg"


def test_binary_errors(tmp_path):
    (tmp_path / "cameras.bin").write_bytes(struct.pack("<Q", 1) + struct.pack("<iiQQ", 1, 42, 1, 1))
    with pytest.raises(F.FormatError):
        F.read_cameras_bin(str(tmp_path / "cameras.bin"))
    (tmp_path / "cameras.bin").write_bytes(struct.pack("<Q", 2) + struct.pack("<iiQQ", 1, 0, 1, 1))
    with pytest.raises(F.FormatError):
        F.read_cameras_bin(str(tmp_path / "cameras.bin"))
    r = _rec()
    r.images[0].name = "x" * 256
    with pytest.raises(F.FormatError):
        F.write_images_bin(r.images, str(tmp_path / "images.bin"))


EXPECTED_CAMERAS_TXT = (
    "# Camera list with one line of data per camera:\n"
    "#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n"
    "# Supported models: SIMPLE_PINHOLE, PINHOLE, SIMPLE_RADIAL, RADIAL, OPENCV, OPENCV_FISHEYE, FULL_OPENCV, FOV, "
    "SIMPLE_RADIAL_FISHEYE, RADIAL_FISHEYE, THIN_PRISM_FISHEYE\n"
    "1 SIMPLE_RADIAL 640 480 500 320 240 0.01\n"
    "2 PINHOLE 800 600 600 610 400.5 299.5\n")
EXPECTED_IMAGES_TXT = (
    "# Image list with two lines of data per image:\n"
    "#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n"
    "#   POINTS2D[] as (X, Y, POINT3D_ID)\n"
    "1 1 0 0 0 0 0 0 1 a.png\n"
    "10.5 20.25 7 30 40 -1\n"
    "2 0.5 0.5 0.5 0.5 1 -2 0.5 2 dir/b c.jpg\n"
    "\n")
EXPECTED_POINTS_TXT = (
    "# 3D point list with one line of data per point:\n"
    "#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n"
    "7 0.1 -0.2 3 255 0 17 0.75 1 0 2 5\n")


def test_text_exact_and_round_trip(tmp_path):
    F.write_txt_dir(_rec(), str(tmp_path))
    assert (tmp_path / "cameras.txt").read_text() == EXPECTED_CAMERAS_TXT
    assert (tmp_path / "images.txt").read_text() == EXPECTED_IMAGES_TXT
    assert (tmp_path / "points3D.txt").read_text() == EXPECTED_POINTS_TXT
    _same(F.read_txt_dir(str(tmp_path)), _rec())
    assert F.read_dir_any(str(tmp_path)).cameras[1].params == [600.0, 610.0, 400.5, 299.5]


def test_text_reader_quirks(tmp_path):
    (tmp_path / "cameras.txt").write_text("# c\n\n  # indented comment\n3 pinhole 10 20 1 2 3 4 99\n1 SIMPLE_PINHOLE 5\n"
                                          "2 SIMPLE_PINHOLE -1 4 7 8 9\n")
    # image 5's line has 9 tokens: it and the line after it are skipped together
    (tmp_path / "images.txt").write_text("9 1 0 0 0 1 2 3 2 my image.png\r\n1 2 3 4 5\r\n"
                                         "5 1 0 0 0 0 0 0 3\n1 1 1\n"
                                         "4 1 0 0 0 0 0 0 3 z.png\n")
    (tmp_path / "points3D.txt").write_text("2 0 0 0 300 -1 5 0.5 9 1 9\n1 1 1 1 1 1 1 1\n")
    r = F.read_txt_dir(str(tmp_path))
    assert [c.camera_id for c in r.cameras] == [2, 3]                     # sorted; 5-token line skipped
    assert r.cameras[0].width == (1 << 64) - 1                           # -1 as u64
    assert r.cameras[1].model_id == F.PINHOLE and r.cameras[1].params == [1.0, 2.0, 3.0, 4.0]
    assert [i.image_id for i in r.images] == [4, 9]
    im9 = r.images[1]
    assert im9.name == "my image.png" and im9.point3d_ids.tolist() == [3]   # incomplete triple dropped
    assert im9.points2d.tolist() == [[1.0, 2.0]]
    assert r.images[0].name == "z.png" and len(r.images[0].point3d_ids) == 0
    p2 = [p for p in r.points3d if p.point3d_id == 2][0]
    assert tuple(p2.rgb) == (300 & 255, 255, 5)                           # `as u8` wraps
    assert p2.track.tolist() == [[9, 1]]                                  # dangling 9 ignored
    (tmp_path / "cameras.txt").write_text("1 NOPE 1 1 1\n")
    with pytest.raises(F.FormatError):
        F.read_txt_dir(str(tmp_path))


# --------------------------------------------------------------------------
# reconstruction -> views
# --------------------------------------------------------------------------

def test_cam_from_rec_models_and_working_size():
    c = lambda mid, p: F.Camera(1, mid, 1000, 500, p)
    assert F.cam_from_rec(c(F.SIMPLE_PINHOLE, [100, 50, 25]), 0.5, 0.5) == (50.0, 25.0, 12.5, 0.0)
    assert F.cam_from_rec(c(F.SIMPLE_RADIAL, [100, 50, 25, 0.1]), 1.0, 1.0) == (100.0, 50.0, 25.0, 0.1)
    assert F.cam_from_rec(c(F.RADIAL, [100, 0.1, 50, 25, 0.2]), 1.0, 1.0) == (100.0, 50.0, 25.0, 0.0)
    assert F.cam_from_rec(c(F.PINHOLE, [100, 110, 50, 25]), 0.5, 0.25) == (105.0 * 0.375, 25.0, 6.25, 0.0)
    assert F.cam_from_rec(c(F.OPENCV, [100, 110, 50, 25, 1, 2, 3, 4]), 1, 1) == (105.0, 50.0, 25.0, 0.0)
    assert F.cam_from_rec(c(F.FOV, [100, 110, 50, 25, 0.3]), 1, 1) == (100.0, 50.0, 25.0, 0.0)
    assert F.cam_from_rec(c(F.SIMPLE_PINHOLE, [100]), 1, 1) == (100.0, 0.0, 0.0, 0.0)   # missing params -> 0
    assert F.working_size(3000, 2000, 1600) == (1600, 1067)       # 2000 * 1600 / 3000 = 1066.67
    assert F.working_size(640, 480, 1600) == (640, 480)           # never upscaled
    assert F.working_size(4000, 10, 1000) == (1000, 16)           # at least 16
    assert F.working_size(2000, 1001, 1000) == (1000, 501)        # 500.5 rounds away from zero


def test_quat_to_rot():
    assert np.array_equal(F.quat_to_rot((1, 0, 0, 0)), np.eye(3))
    s = math.sqrt(0.5)
    R = F.quat_to_rot((s, 0, 0, s))                                # 90 degrees about z
    assert np.allclose(R, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-15)
    assert np.allclose(F.quat_to_rot((2, 0, 0, 0)), np.eye(3))     # norm > 1: renormalised
    R = F.quat_to_rot((0.5, 0, 0, 0))                              # norm < 1: used as is (reference quirk)
    assert np.array_equal(R, np.eye(3))                            # 1 - 2 * 0 on the diagonal
    R = F.quat_to_rot((0.5, 0.5, 0, 0))
    assert R[1, 2] == 2.0 * (0.0 - 0.5 * 0.5)


def test_views_from_reconstruction():
    rec = _rec()
    rec.images[1].camera_id = 1
    images = {"a.png": (np.zeros((480, 640), np.uint8), None),
              "dir/b c.jpg": (np.full((480, 640), 7, np.uint8), np.zeros((480, 640, 3), np.uint8))}
    views, xyz, tracks = F.views_from_reconstruction(rec, images, 1600)
    assert [v.name for v in views] == ["a.png", "dir/b c.jpg"] and [v.image_id for v in views] == [1, 2]
    assert (views[0].cam.f, views[0].cam.cx, views[0].cam.cy, views[0].cam.k) == (500.0, 320.0, 240.0, 0.01)
    assert np.array_equal(views[1].pose.t, [1.0, -2.0, 0.5])
    assert np.allclose(views[1].pose.r, F.quat_to_rot((0.5, 0.5, 0.5, 0.5)))
    assert xyz.tolist() == [[0.1, -0.2, 3.0]] and tracks == [[0, 1]]
    with pytest.raises(F.FormatError):
        F.views_from_reconstruction(rec, images, 320)              # pixels not at the 320x240 working size
    with pytest.raises(F.FormatError):
        F.views_from_reconstruction(rec, {"a.png": images["a.png"]}, 1600)


# --------------------------------------------------------------------------
# dense_web.json
# --------------------------------------------------------------------------

def _cloud(xyz, rgb=None):
    from rustcolmap_amd.mvs import POINT_DTYPE
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    p = np.zeros(len(xyz), POINT_DTYPE)
    p["x"], p["y"], p["z"] = xyz.T
    if rgb is not None:
        p["rgb"] = np.asarray(rgb)
    return p


def test_dense_web_json_closed_form():
    # 20 points on the unit circle plus one at x = 1000: centroid (1000/21, 0,
    # 0); the circle lies 46.6-48.6 from it, the far point 952.4.  Sorted,
    # med = ds[10] and p95 = ds[int(21 * 0.95)] = ds[19] are both circle
    # distances, so r_keep = med + 6 * max(|p95 - med|, med / 2) ~ 4 med ~ 190:
    # the far point is dropped, the circle kept
    ang = np.linspace(0, 2 * np.pi, 20, endpoint=False)
    xyz = np.stack([np.cos(ang), np.sin(ang), np.zeros(20)], 1).tolist() + [[1000.0, 0.0, 0.0]]
    js = json.loads(F.dense_web_json(_cloud(xyz)))
    assert js["n"] == 20 and [1000.0, 0.0, 0.0] not in js["pos"]
    # with 2 far points of 22, p95 = ds[20] is a far one: the radius grows to
    # med + 6 (952 - med) and keeps everything
    js = json.loads(F.dense_web_json(_cloud(xyz + [[1000.0, 1.0, 0.0]])))
    assert js["n"] == 22
    # a tight cluster of 100 points and one point far out: p95 sits in the
    # cluster, so the keep radius is med + 6 * max(|p95 - med|, med / 2)
    rng = np.random.default_rng(0)
    c = rng.normal(size=(100, 3)) * 0.01 + 5.0
    xyz = np.concatenate([c, [[500.0, 5.0, 5.0]]])
    s = F.dense_web_json(_cloud(xyz, np.tile([1, 2, 3], (101, 1))))
    js = json.loads(s)
    assert js["n"] == 100 and [500.0, 5.0, 5.0] not in js["pos"]
    assert s.count("[1, 2,3]") == 100 and '"col":[[1, 2,3],' in s
    x0 = xyz[0]
    assert s.startswith('{"n":100,"pos":[[%.4f,%.4f,%.4f],' % tuple(x0))
    # decimation: every ceil(kept / cap)-th point
    s = F.dense_web_json(_cloud(c), cap=30)
    js = json.loads(s)
    assert js["n"] == 25 and js["pos"][1] == [round(v, 4) for v in c[4]]
    assert F.dense_web_json(_cloud(np.zeros((0, 3)))) is None


def test_dense_web_json_c_abi_identical(tmp_path):
    from rustcolmap_amd import _lib
    try:
        lib = _lib.load()
    except ImportError:
        pytest.fail("HIP library not built")
    rng = np.random.default_rng(3)
    xyz = np.concatenate([rng.normal(size=(5000, 3)), rng.normal(size=(40, 3)) * 200])
    p = _cloud(xyz, rng.integers(0, 256, size=(len(xyz), 3)))
    path = str(tmp_path / "w.json")
    _lib.check(lib.mvs_write_dense_web_json(path.encode(), p.ctypes.data_as(C.c_void_p), len(p)),
               "mvs_write_dense_web_json")
    with open(path) as f:
        assert f.read() == F.dense_web_json(p)
    empty = str(tmp_path / "e.json")
    _lib.check(lib.mvs_write_dense_web_json(empty.encode(), None, 0), "mvs_write_dense_web_json")
    assert not os.path.exists(empty)


# --------------------------------------------------------------------------
# depth visualisation
# --------------------------------------------------------------------------

def test_depth_vis_closed_form(tmp_path):
    d = np.zeros((10, 10), np.float32)
    d.ravel()[:100] = np.arange(1, 101, dtype=np.float32)     # 100 valid depths 1..100
    d[0, 0] = 0.0                                             # one invalid: 99 valid 2..100
    img = F.depth_vis(d)
    lo, hi = np.float32(2 + int(99 * 0.02)), np.float32(2 + int(99 * 0.98) - 1)   # 3 and 98
    assert img[0, 0] == 0
    assert img.ravel()[2] == 255                              # depth 3 = lo: nearest -> brightest
    assert img.ravel()[97] == 0 and img.ravel()[99] == 0       # depth 98 = hi and beyond -> 0
    z = np.float32(50.0)
    t = (z - lo) / (hi - lo)
    assert img.ravel()[49] == int(np.float32(255.0) * (np.float32(1.0) - np.float32(t)))
    few = np.where(np.arange(100).reshape(10, 10) < 10, np.float32(0.5), np.float32(0)).astype(np.float32)
    assert set(F.depth_vis(few).ravel().tolist()) == {0, int(np.float32(255) * np.float32(0.5))}   # range [0, 1]
    path = F.save_depth_vis(str(tmp_path), "imgs/view_07.jpg", d)
    assert os.path.basename(path) == "view_07_depth.png"
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n" and data[12:16] == b"IHDR"
    assert struct.unpack(">IIBB", data[16:26]) == (10, 10, 8, 0)
    assert zlib.crc32(data[12:29]) & 0xFFFFFFFF == struct.unpack(">I", data[29:33])[0]
    assert np.array_equal(F.read_png_gray8(data), img)


# --------------------------------------------------------------------------
# persisted depth maps
# --------------------------------------------------------------------------

def test_depth_map_container(tmp_path):
    rng = np.random.default_rng(1)
    d = rng.random((7, 9)).astype(np.float32)
    c = rng.random((7, 9)).astype(np.float32)
    n = rng.random((7, 9, 3)).astype(np.float32)
    p = str(tmp_path / "v.depth")
    F.write_depth_map(p, d, c, n)
    d2, c2, n2 = F.read_depth_map(p)
    assert np.array_equal(d, d2) and np.array_equal(c, c2) and np.array_equal(n, n2)
    b = open(p, "rb").read()
    assert b[:8] == b"MVSDEPTH" and struct.unpack("<IIII", b[8:24]) == (1, 9, 7, 3) and len(b) == 24 + 63 * 20
    F.write_depth_map(p, d)
    assert F.read_depth_map(p)[1] is None
    open(p, "wb").write(b"garbage!" + b[8:])
    with pytest.raises(F.FormatError):
        F.read_depth_map(p)


def test_rust_f64_debug():
    """`{:?}` of f64 as Rust prints it (core::fmt::float: positional for
    1e-4 <= |x| < 1e16, exponential outside, shortest round-trip digits)."""
    from rustcolmap_amd.formats import rust_f64_debug as d
    cases = [(1.0, "1.0"), (10.0, "10.0"), (0.1, "0.1"), (-2.5, "-2.5"), (0.0, "0.0"), (-0.0, "-0.0"),
             (1e-4, "0.0001"), (9.99e-5, "9.99e-5"), (1e-5, "1e-5"), (1e16, "1e16"), (1.5e16, "1.5e16"),
             (9999999999999998.0, "9999999999999998.0"), (123456.789, "123456.789"), (1.5e300, "1.5e300"),
             (2.2250738585072014e-308, "2.2250738585072014e-308"), (0.30000000000000004, "0.30000000000000004"),
             (12345678901234567.0, "1.2345678901234568e16"), (float("inf"), "inf"), (float("-inf"), "-inf"),
             (4.0 * 0.8, "3.2"), (1.25 * 10, "12.5")]
    for x, want in cases:
        assert d(x) == want, (x, d(x), want)
    assert d(float("nan")) == "NaN"


def test_library_rust_f64_debug_matches():
    """The library's own `{:?}` formatter (host code of the log lines) gives
    the same strings as the restatement above, on edge cases and on random
    doubles over the whole exponent range."""
    import ctypes as C
    import numpy as np
    from rustcolmap_amd import _lib
    from rustcolmap_amd.formats import rust_f64_debug
    lib = _lib.load()
    buf = C.create_string_buffer(64)
    rng = np.random.default_rng(4)
    xs = [1.0, 10.0, 0.1, -2.5, 0.0, -0.0, 1e-4, 9.99e-5, 1e-5, 1e16, 1.5e16, 9999999999999998.0, 5e-324,
          1.7976931348623157e308, 0.30000000000000004, 12345678901234567.0, 2.0 ** 53, 1.0 / 3.0]
    xs += list(rng.uniform(0.1, 100.0, 2000)) + list(np.exp(rng.uniform(-700, 700, 2000)) * rng.choice([-1, 1], 2000))
    xs += list(np.round(rng.uniform(0, 50, 500), 2))
    for x in xs:
        _lib.check(lib.mvs_debug_rust_f64(float(x), buf, 64), "mvs_debug_rust_f64")
        assert buf.value.decode() == rust_f64_debug(float(x)), (repr(x), buf.value, rust_f64_debug(float(x)))
