"""On-disk formats on either side of the dense path.

* COLMAP sparse models (`cameras`, `images`, `points3D`) in `.bin` and `.txt`,
  read and written byte for byte as the reference's `colmap-io` crate does
  (colmap-io/src/lib.rs:14-584): little-endian binary records, and text whose
  floats are printed the way Rust's `{}` prints an f64 (shortest round-trip
  digits, never an exponent).
* `views_from_reconstruction`: the reference's `cam_from_rec` and the
  working-resolution arithmetic of `load_views` (mvs/src/lib.rs:103-117,
  144-218) for decoded images, resized on the device when they come at
  their native size.
* `dense_web.json` (mvs/src/lib.rs:437-485), the decimated display copy of the
  fused cloud, and `save_depth_vis` (mvs/src/fuse.rs:241-269), the grayscale
  depth visualisation as an 8-bit PNG.
* Depth maps persisted per view (`write_depth_map` / `read_depth_map`) so that
  fusion can be rerun without stereo; the reference keeps them in memory only,
  so this container format is this library's own.
"""
from __future__ import annotations

import math
import os
import struct
import zlib
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# --------------------------------------------------------------------------
# camera models (colmap-io/src/lib.rs:18-56): ids and parameter counts
# --------------------------------------------------------------------------

SIMPLE_PINHOLE, PINHOLE, SIMPLE_RADIAL, RADIAL, OPENCV, OPENCV_FISHEYE = 0, 1, 2, 3, 4, 5
FULL_OPENCV, FOV, SIMPLE_RADIAL_FISHEYE, RADIAL_FISHEYE, THIN_PRISM_FISHEYE = 6, 7, 8, 9, 10

MODELS: Tuple[Tuple[int, str, int], ...] = (
    (SIMPLE_PINHOLE, "SIMPLE_PINHOLE", 3), (PINHOLE, "PINHOLE", 4), (SIMPLE_RADIAL, "SIMPLE_RADIAL", 4),
    (RADIAL, "RADIAL", 5), (OPENCV, "OPENCV", 8), (OPENCV_FISHEYE, "OPENCV_FISHEYE", 8),
    (FULL_OPENCV, "FULL_OPENCV", 12), (FOV, "FOV", 5), (SIMPLE_RADIAL_FISHEYE, "SIMPLE_RADIAL_FISHEYE", 4),
    (RADIAL_FISHEYE, "RADIAL_FISHEYE", 5), (THIN_PRISM_FISHEYE, "THIN_PRISM_FISHEYE", 12),
)


def model_by_id(mid: int):
    for m in MODELS:
        if m[0] == mid:
            return m
    return None


def model_by_name(name: str):
    for m in MODELS:
        if m[1].lower() == name.lower():
            return m
    return None


class FormatError(ValueError):
    """Malformed or unsupported file contents (the reference's InvalidData)."""


# --------------------------------------------------------------------------
# data model (colmap-io/src/lib.rs:62-125)
# --------------------------------------------------------------------------

@dataclass
class Camera:
    camera_id: int
    model_id: int
    width: int
    height: int
    params: List[float]


@dataclass
class Image:
    """World->camera pose x_cam = R(qvec) x_world + tvec, qvec = (qw, qx, qy, qz).
    points2d: (n, 2) float64 pixel positions; point3d_ids: (n,) int64, -1 = none."""
    image_id: int
    qvec: Tuple[float, float, float, float]
    tvec: Tuple[float, float, float]
    camera_id: int
    name: str
    points2d: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float64))
    point3d_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))


@dataclass
class Point3D:
    """track: (n, 2) int64 rows of (image_id, point2d_idx)."""
    point3d_id: int
    xyz: Tuple[float, float, float]
    rgb: Tuple[int, int, int]
    error: float
    track: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))


@dataclass
class Reconstruction:
    cameras: List[Camera] = field(default_factory=list)
    images: List[Image] = field(default_factory=list)
    points3d: List[Point3D] = field(default_factory=list)

    def camera(self, cid: int) -> Optional[Camera]:
        for c in self.cameras:
            if c.camera_id == cid:
                return c
        return None

    def image_by_name(self, name: str) -> Optional[Image]:
        for i in self.images:
            if i.name == name:
                return i
        return None

    def sort_all(self):
        """Stable sorts by id (Rust's sort_by_key)."""
        self.cameras.sort(key=lambda c: c.camera_id)
        self.images.sort(key=lambda i: i.image_id)
        self.points3d.sort(key=lambda p: p.point3d_id)


# --------------------------------------------------------------------------
# Rust number formatting
# --------------------------------------------------------------------------

def rust_f64(x: float) -> str:
    """`format!("{}", x)` for an f64: the shortest digits that round-trip
    (the same digits as Python's repr), written out without an exponent;
    integral values carry no fraction ("1", "-0"); NaN, inf, -inf."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    r = repr(abs(x))
    if "e" in r or "E" in r:
        mant, ex = r.lower().split("e")
        exp = int(ex)
    else:
        mant, exp = r, 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    digits = (ip + fp).lstrip("0")
    # position of the decimal point relative to the start of `digits`
    lead_zeros = len(ip + fp) - len((ip + fp).lstrip("0"))
    point = len(ip) + exp - lead_zeros
    digits = digits.rstrip("0")
    if not digits:
        return sign + "0"
    if point <= 0:
        s = "0." + "0" * (-point) + digits
    elif point >= len(digits):
        s = digits + "0" * (point - len(digits))
    else:
        s = digits[:point] + "." + digits[point:]
    return sign + s


def rust_f64_debug(x: float) -> str:
    """`format!("{:?}", x)` for an f64: the shortest round-trip digits in
    positional notation with at least one fraction digit ("1.0", "0.0001")
    for 1e-4 <= |x| < 1e16 and zero, otherwise as <d>[.<ddd>]e<exp>
    ("1e-5", "1.5e16"); NaN, inf, -inf."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    m, e = ("%r" % abs(x)).lower().split("e") if "e" in repr(abs(x)).lower() else (repr(abs(x)), "0")
    ip, _, fp = m.partition(".")
    digits = (ip + fp).lstrip("0")
    lead = len(ip + fp) - len(digits)
    exp10 = len(ip) - 1 + int(e) - lead          # power of ten of the first significant digit
    digits = digits.rstrip("0") or "0"
    if abs(x) < 1e-4 or abs(x) >= 1e16:
        return sign + digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e%d" % exp10
    s = rust_f64(abs(x))
    return sign + (s if "." in s else s + ".0")


def _wrap(v: int, bits: int, signed: bool) -> int:
    """Rust `as` between integer types: keep the low `bits` bits."""
    v &= (1 << bits) - 1
    if signed and v >= 1 << (bits - 1):
        v -= 1 << bits
    return v


# --------------------------------------------------------------------------
# binary (colmap-io/src/lib.rs:244-394)
# --------------------------------------------------------------------------

class _Reader:
    def __init__(self, path: str):
        self.path = path
        with open(path, "rb") as f:
            self.b = f.read()
        self.o = 0

    def take(self, n: int) -> bytes:
        if self.o + n > len(self.b):
            raise FormatError("%s: unexpected end of file" % self.path)
        v = self.b[self.o:self.o + n]
        self.o += n
        return v

    def unpack(self, fmt: str):
        return struct.unpack("<" + fmt, self.take(struct.calcsize("<" + fmt)))

    def array(self, dtype, n: int) -> np.ndarray:
        dt = np.dtype(dtype).newbyteorder("<")
        return np.frombuffer(self.take(dt.itemsize * n), dt).astype(dt.newbyteorder("="))


def read_cameras_bin(path: str) -> List[Camera]:
    r = _Reader(path)
    (n,) = r.unpack("Q")
    out = []
    for _ in range(n):
        cid, mid, w, h = r.unpack("iiQQ")
        m = model_by_id(mid)
        if m is None:
            raise FormatError("unknown camera model id %d" % mid)
        out.append(Camera(cid, mid, w, h, list(r.unpack("%dd" % m[2]))))
    return out


def read_images_bin(path: str) -> List[Image]:
    r = _Reader(path)
    (n,) = r.unpack("Q")
    out = []
    for _ in range(n):
        (iid,) = r.unpack("i")
        q = r.unpack("4d")
        t = r.unpack("3d")
        (cid,) = r.unpack("i")
        (ln,) = r.unpack("B")
        try:
            name = r.take(ln).decode("utf-8")
        except UnicodeDecodeError as e:
            raise FormatError("bad utf8 name: %s" % e)
        (npt,) = r.unpack("Q")
        rec = r.array([("x", "f8"), ("y", "f8"), ("id", "i8")], npt)
        xy = np.stack([rec["x"], rec["y"]], axis=1) if npt else np.zeros((0, 2), np.float64)
        out.append(Image(iid, q, t, cid, name, xy.astype(np.float64), rec["id"].astype(np.int64)))
    return out


def read_points3d_bin(path: str) -> List[Point3D]:
    r = _Reader(path)
    (n,) = r.unpack("Q")
    out = []
    for _ in range(n):
        pid, x, y, z, cr, cg, cb, err, tl = r.unpack("qdddBBBdQ")
        tr = r.array("i4", 2 * tl).reshape(-1, 2).astype(np.int64)
        tr[:, 1] &= 0xFFFFFFFF                   # point2d_idx read as i32, kept as u32
        out.append(Point3D(pid, (x, y, z), (cr, cg, cb), err, tr))
    return out


def _open_w(path: str):
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    return open(path, "wb")


def write_cameras_bin(cameras: Sequence[Camera], path: str):
    with _open_w(path) as f:
        f.write(struct.pack("<Q", len(cameras)))
        for c in cameras:
            f.write(struct.pack("<iiQQ", c.camera_id, c.model_id, c.width, c.height))
            f.write(struct.pack("<%dd" % len(c.params), *c.params))


def write_images_bin(images: Sequence[Image], path: str):
    with _open_w(path) as f:
        f.write(struct.pack("<Q", len(images)))
        for im in images:
            nb = im.name.encode("utf-8")
            if len(nb) > 255:
                raise FormatError("image name longer than 255 bytes: %d" % len(nb))
            f.write(struct.pack("<i4d3di", im.image_id, *im.qvec, *im.tvec, im.camera_id))
            f.write(struct.pack("<B", len(nb)) + nb)
            n = len(im.point3d_ids)
            f.write(struct.pack("<Q", n))
            rec = np.zeros(n, [("x", "<f8"), ("y", "<f8"), ("id", "<i8")])
            if n:
                rec["x"], rec["y"], rec["id"] = im.points2d[:, 0], im.points2d[:, 1], im.point3d_ids
            f.write(rec.tobytes())


def write_points3d_bin(points: Sequence[Point3D], path: str):
    with _open_w(path) as f:
        f.write(struct.pack("<Q", len(points)))
        for p in points:
            f.write(struct.pack("<q3d3BdQ", p.point3d_id, *p.xyz, *p.rgb, p.error, len(p.track)))
            tr = np.asarray(p.track, np.int64).reshape(-1, 2)
            f.write(np.stack([tr[:, 0].astype("<i4"), (tr[:, 1] & 0xFFFFFFFF).astype(np.uint32).view("<i4")],
                             axis=1).astype("<i4").tobytes())


def write_bin_dir(rec: Reconstruction, d: str):
    write_cameras_bin(rec.cameras, os.path.join(d, "cameras.bin"))
    write_images_bin(rec.images, os.path.join(d, "images.bin"))
    write_points3d_bin(rec.points3d, os.path.join(d, "points3D.bin"))


def read_bin_dir(d: str) -> Reconstruction:
    return Reconstruction(read_cameras_bin(os.path.join(d, "cameras.bin")),
                          read_images_bin(os.path.join(d, "images.bin")),
                          read_points3d_bin(os.path.join(d, "points3D.bin")))


# --------------------------------------------------------------------------
# text (colmap-io/src/lib.rs:400-575)
# --------------------------------------------------------------------------

def write_txt_dir(rec: Reconstruction, d: str):
    os.makedirs(d, exist_ok=True)
    s = ["# Camera list with one line of data per camera:\n",
         "#   CAMERA_ID, MODEL, WIDTH, HEIGHT, PARAMS[]\n",
         "# Supported models: %s\n" % ", ".join(m[1] for m in MODELS)]
    for c in rec.cameras:
        m = model_by_id(c.model_id)
        s.append("%d %s %d %d" % (c.camera_id, m[1] if m else "UNKNOWN", c.width, c.height))
        s.extend(" " + rust_f64(p) for p in c.params)
        s.append("\n")
    with open(os.path.join(d, "cameras.txt"), "w", encoding="utf-8", newline="") as f:
        f.write("".join(s))
    s = ["# Image list with two lines of data per image:\n",
         "#   IMAGE_ID, QW, QX, QY, QZ, TX, TY, TZ, CAMERA_ID, NAME\n",
         "#   POINTS2D[] as (X, Y, POINT3D_ID)\n"]
    for im in rec.images:
        s.append(" ".join([str(im.image_id)] + [rust_f64(v) for v in tuple(im.qvec) + tuple(im.tvec)]
                          + [str(im.camera_id), im.name]) + "\n")
        s.append(" ".join("%s %s %d" % (rust_f64(x), rust_f64(y), int(i))
                          for (x, y), i in zip(im.points2d, im.point3d_ids)) + "\n")
    with open(os.path.join(d, "images.txt"), "w", encoding="utf-8", newline="") as f:
        f.write("".join(s))
    s = ["# 3D point list with one line of data per point:\n",
         "#   POINT3D_ID, X, Y, Z, R, G, B, ERROR, TRACK[] as (IMAGE_ID, POINT2D_IDX)\n"]
    for p in rec.points3d:
        s.append("%d %s %s %s %d %d %d %s" % (p.point3d_id, rust_f64(p.xyz[0]), rust_f64(p.xyz[1]),
                                              rust_f64(p.xyz[2]), p.rgb[0], p.rgb[1], p.rgb[2], rust_f64(p.error)))
        s.extend(" %d %d" % (int(a), int(b)) for a, b in np.asarray(p.track).reshape(-1, 2))
        s.append("\n")
    with open(os.path.join(d, "points3D.txt"), "w", encoding="utf-8", newline="") as f:
        f.write("".join(s))


def _data_lines(path: str) -> List[str]:
    with open(path, "r", encoding="utf-8", newline="") as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()                                  # BufRead::lines drops the final terminator
    lines = [l[:-1] if l.endswith("\r") else l for l in lines]
    return [l for l in lines if not l.lstrip().startswith("#")]


def _f(tok: str) -> float:
    try:
        return float(tok)
    except ValueError:
        raise FormatError("bad float `%s`" % tok)


def _i(tok: str) -> int:
    try:
        if not tok or not (tok.lstrip("+-").isdigit()) or tok.count("-") + tok.count("+") > 1:
            raise ValueError
        return int(tok)
    except ValueError:
        raise FormatError("bad int `%s`" % tok)


def read_txt_dir(d: str) -> Reconstruction:
    cams = []
    for line in _data_lines(os.path.join(d, "cameras.txt")):
        t = line.split()
        if len(t) < 5:
            continue
        m = model_by_name(t[1])
        if m is None:
            raise FormatError("unknown model `%s`" % t[1])
        cams.append(Camera(_wrap(_i(t[0]), 32, True), m[0], _wrap(_i(t[2]), 64, False),
                           _wrap(_i(t[3]), 64, False), [_f(x) for x in t[4:4 + m[2]]]))
    data = _data_lines(os.path.join(d, "images.txt"))
    imgs = []
    k = 0
    while k < len(data):
        l1 = data[k]
        l2 = data[k + 1] if k + 1 < len(data) else ""
        k += 2
        t = l1.split()
        if len(t) < 10:
            continue
        toks = l2.split()
        trip = [toks[j:j + 3] for j in range(0, len(toks), 3)]
        trip = [c for c in trip if len(c) == 3]
        xy = np.array([[_f(a), _f(b)] for a, b, _ in trip], np.float64).reshape(-1, 2)
        ids = np.array([_i(c) for _, _, c in trip], np.int64)
        imgs.append(Image(_wrap(_i(t[0]), 32, True), tuple(_f(x) for x in t[1:5]), tuple(_f(x) for x in t[5:8]),
                          _wrap(_i(t[8]), 32, True), " ".join(t[9:]), xy, ids))
    pts = []
    for line in _data_lines(os.path.join(d, "points3D.txt")):
        t = line.split()
        if len(t) < 8:
            continue
        pairs = [t[8 + j:10 + j] for j in range(0, len(t) - 8, 2)]
        pairs = [p for p in pairs if len(p) == 2]
        tr = np.array([[_wrap(_i(a), 32, True), _wrap(_i(b), 32, False)] for a, b in pairs],
                      np.int64).reshape(-1, 2)
        pts.append(Point3D(_i(t[0]), (_f(t[1]), _f(t[2]), _f(t[3])),
                           (_wrap(_i(t[4]), 8, False), _wrap(_i(t[5]), 8, False), _wrap(_i(t[6]), 8, False)),
                           _f(t[7]), tr))
    rec = Reconstruction(cams, imgs, pts)
    rec.sort_all()
    return rec


def read_dir_any(d: str) -> Reconstruction:
    """Binary when `cameras.bin` exists, else text (colmap-io/src/lib.rs:578-584)."""
    return read_bin_dir(d) if os.path.exists(os.path.join(d, "cameras.bin")) else read_txt_dir(d)


# --------------------------------------------------------------------------
# reconstruction -> stereo inputs (mvs/src/lib.rs:103-117, 144-218, 224-243)
# --------------------------------------------------------------------------

def cam_from_rec(c: Camera, sx: float, sy: float):
    """(f, cx, cy, k) at working resolution (mvs/src/lib.rs:103-117)."""
    g = lambda i: float(c.params[i]) if i < len(c.params) else 0.0
    mid = c.model_id
    if mid == SIMPLE_PINHOLE:
        f, cx, cy, k = g(0), g(1), g(2), 0.0
    elif mid == SIMPLE_RADIAL:
        f, cx, cy, k = g(0), g(1), g(2), g(3)
    elif mid == RADIAL:
        f, cx, cy, k = g(0), g(2), g(3), 0.0
    elif mid in (PINHOLE, OPENCV):
        f, cx, cy, k = 0.5 * (g(0) + g(1)), g(2), g(3), 0.0
    else:
        f, cx, cy, k = g(0), g(2), g(3), 0.0
    s = 0.5 * (sx + sy)
    return f * s, cx * sx, cy * sy, k


def working_size(width: int, height: int, max_dim: int) -> Tuple[int, int]:
    """load_views' resize target (mvs/src/lib.rs:154-157): round, at least 16."""
    s = min(float(max_dim) / float(max(width, height)), 1.0)
    rnd = lambda v: int(math.floor(v + 0.5))            # f64::round on non-negative values
    return max(rnd(width * s), 16), max(rnd(height * s), 16)


def quat_to_rot(q) -> np.ndarray:
    """COLMAP (qw, qx, qy, qz) -> R, f64 in the reference's operation order
    (sfm-lite/src/geom.rs:79-93); like the reference it renormalises only
    when |q|^2 - 1 > 1e-12 (a quaternion of norm below 1 is used as is)."""
    w, x, y, z = (float(v) for v in q)
    n = w * w + x * x + y * y + z * z
    if abs(n) - 1.0 > 1e-12:
        s = math.sqrt(n)
        w, x, y, z = w / s, x / s, y / s, z / s
    return np.array([[1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)],
                     [2.0 * (x * y + w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - w * x)],
                     [2.0 * (x * z - w * y), 2.0 * (y * z + w * x), 1.0 - 2.0 * (x * x + y * y)]], np.float64)


def views_from_reconstruction(rec: Reconstruction, images: Dict[str, Tuple[Optional[np.ndarray], Optional[np.ndarray]]],
                              max_dim: int, ctx=None):
    """ViewData for every image of `rec` (in `rec.images` order).
    images[name] = (gray, rgb): either at the working size load_views
    produces (gray required, rgb optional), or one of them decoded at the
    camera's native size, which is then resized and converted on the device
    as load_views does (mvs.resize_image; needs `ctx`).  Returns (views,
    xyz (n, 3) float64, tracks as lists of view indices).  Raises
    FormatError for a missing camera or image or an unexpected size."""
    from .mvs import Cam, Pose, ViewData, resize_image
    views = []
    for im in rec.images:
        c = rec.camera(im.camera_id)
        if c is None:
            raise FormatError("image %s references missing camera %d" % (im.name, im.camera_id))
        mw, mh = working_size(c.width, c.height, max_dim)
        f, cx, cy, k = cam_from_rec(c, mw / float(c.width), mh / float(c.height))
        if im.name not in images:
            raise FormatError("no pixels for image %s" % im.name)
        gray, rgb = images[im.name]
        src = rgb if rgb is not None else gray
        if src is None:
            raise FormatError("no pixels for image %s" % im.name)
        src = np.ascontiguousarray(src, np.uint8)
        if src.shape[:2] == (c.height, c.width) and (c.height, c.width) != (mh, mw):
            if ctx is None:
                raise FormatError("image %s is at native size; resizing it needs a context" % im.name)
            gray, rgb = resize_image(ctx, src, mw, mh)
        if gray is None:
            raise FormatError("image %s: luma missing at the working size" % im.name)
        gray = np.ascontiguousarray(gray, np.uint8)
        if gray.shape != (mh, mw):
            raise FormatError("image %s is %s, working size is %dx%d" % (im.name, gray.shape[::-1], mw, mh))
        if rgb is not None:
            rgb = np.ascontiguousarray(rgb, np.uint8).reshape(mh, mw, 3)
        views.append(ViewData(w=mw, h=mh, cam=Cam(f, cx, cy, k),
                              pose=Pose(quat_to_rot(im.qvec), np.array(im.tvec, np.float64)),
                              gray=gray, rgb=rgb, name=im.name, image_id=im.image_id))
    idx = {im.image_id: i for i, im in enumerate(rec.images)}
    xyz = np.array([p.xyz for p in rec.points3d], np.float64).reshape(-1, 3)
    tracks = [[idx[int(t)] for t in np.asarray(p.track).reshape(-1, 2)[:, 0] if int(t) in idx]
              for p in rec.points3d]
    return views, xyz, tracks


# --------------------------------------------------------------------------
# dense_web.json (mvs/src/lib.rs:437-485) and depth visualisation
# --------------------------------------------------------------------------

WEB_CAP = 400_000


def dense_web_json(points: np.ndarray, cap: int = WEB_CAP) -> Optional[str]:
    """Decimated display copy of a fused cloud: outliers beyond the robust
    radius dropped, then every ceil(kept / cap)-th point; positions with 4
    decimals.  `points` is an array of mvs.POINT_DTYPE (x, y, z, rgb) in the
    cloud's order.  None for an empty cloud (no file is written then)."""
    n = len(points)
    if n == 0:
        return None
    x = np.asarray(points["x"], np.float64)
    y = np.asarray(points["y"], np.float64)
    z = np.asarray(points["z"], np.float64)
    # f64 sums in point order, as the reference's iterator sums
    sx = sy = sz = 0.0
    for a, b, c in zip(x.tolist(), y.tolist(), z.tolist()):
        sx += a
        sy += b
        sz += c
    cx, cy, cz = sx / n, sy / n, sz / n
    dx, dy, dz = x - cx, y - cy, z - cz
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    ds = np.sort(dist)
    med = float(ds[n // 2])
    p95 = float(ds[min(int(n * 0.95), n - 1)])
    r_keep = med + 6.0 * max(abs(p95 - med), max(med, 1e-9) * 0.5)
    kept = np.flatnonzero(dist <= r_keep)
    stride = -(-len(kept) // cap)
    web = kept[::stride]
    parts = ['{"n":', str(len(web)), ',"pos":[']
    parts.append(",".join("[%.4f,%.4f,%.4f]" % (x[i], y[i], z[i]) for i in web))
    parts.append('],"col":[')
    rgb = np.asarray(points["rgb"]).reshape(-1, 3)
    parts.append(",".join("[%d, %d,%d]" % tuple(int(v) for v in rgb[i]) for i in web))
    parts.append("]}")
    return "".join(parts)


def depth_vis(depth: np.ndarray) -> np.ndarray:
    """u8 image of save_depth_vis (mvs/src/fuse.rs:241-269): 2nd/98th
    percentile of the valid depths, near = bright, invalid = 0, f32 math."""
    d = np.ascontiguousarray(depth, np.float32)
    ds = np.sort(d[d > 0].ravel())
    if len(ds) >= 32:
        lo = ds[int(len(ds) * 0.02)]
        hi = ds[int(len(ds) * 0.98) - 1]
    else:
        lo, hi = np.float32(0.0), np.float32(1.0)
    rng = np.maximum(np.float32(hi - lo), np.float32(1e-9))
    with np.errstate(invalid="ignore"):
        t = np.clip((d - lo) / rng, np.float32(0.0), np.float32(1.0)).astype(np.float32)
        v = (np.float32(255.0) * (np.float32(1.0) - t)).astype(np.float32)
    out = np.where(d > 0, np.clip(np.floor(v), 0, 255), 0).astype(np.uint8)
    return out


def png_gray8(img: np.ndarray) -> bytes:
    """Minimal 8-bit grayscale PNG (no filtering, zlib level 6)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def read_png_gray8(data: bytes) -> np.ndarray:
    """Decoder for what png_gray8 writes (filter type 0 rows)."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise FormatError("not a PNG")
    o = 8
    w = h = None
    idat = b""
    while o < len(data):
        (ln,) = struct.unpack(">I", data[o:o + 4])
        tag = data[o + 4:o + 8]
        body = data[o + 8:o + 8 + ln]
        if tag == b"IHDR":
            w, h, bd, ct, _, _, _ = struct.unpack(">IIBBBBB", body)
            if bd != 8 or ct != 0:
                raise FormatError("only 8-bit grayscale")
        elif tag == b"IDAT":
            idat += body
        o += 12 + ln
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, w + 1)
    if np.any(rows[:, 0] != 0):
        raise FormatError("filtered rows are not supported")
    return rows[:, 1:].copy()


def save_depth_vis(d: str, name: str, depth: np.ndarray) -> str:
    """<d>/<stem>_depth.png (mvs/src/fuse.rs:241-269)."""
    stem = os.path.splitext(os.path.basename(name))[0] or "view"
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, stem + "_depth.png")
    with open(path, "wb") as f:
        f.write(png_gray8(depth_vis(depth)))
    return path


# --------------------------------------------------------------------------
# persisted depth maps (this library's own container)
# --------------------------------------------------------------------------

DEPTH_MAGIC = b"MVSDEPTH"
DEPTH_VERSION = 1


def write_depth_map(path: str, depth: np.ndarray, cost: Optional[np.ndarray] = None,
                    normals: Optional[np.ndarray] = None):
    """Little-endian: magic, u32 version, u32 width, u32 height, u32 flags
    (bit 0 cost, bit 1 normals), then f32 depth[h][w], f32 cost[h][w],
    f32 normals[h][w][3] as present."""
    depth = np.ascontiguousarray(depth, "<f4")
    h, w = depth.shape
    flags = (1 if cost is not None else 0) | (2 if normals is not None else 0)
    with _open_w(path) as f:
        f.write(DEPTH_MAGIC + struct.pack("<IIII", DEPTH_VERSION, w, h, flags))
        f.write(depth.tobytes())
        if cost is not None:
            f.write(np.ascontiguousarray(cost, "<f4").reshape(h, w).tobytes())
        if normals is not None:
            f.write(np.ascontiguousarray(normals, "<f4").reshape(h, w, 3).tobytes())


def read_depth_map(path: str):
    """(depth, cost or None, normals or None) as written by write_depth_map."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != DEPTH_MAGIC:
        raise FormatError("%s: not a depth map" % path)
    ver, w, h, flags = struct.unpack("<IIII", b[8:24])
    if ver != DEPTH_VERSION:
        raise FormatError("%s: unsupported version %d" % (path, ver))
    o = 24
    need = w * h * 4 * (1 + (1 if flags & 1 else 0) + (3 if flags & 2 else 0))
    if len(b) != o + need:
        raise FormatError("%s: size mismatch" % path)
    depth = np.frombuffer(b, "<f4", w * h, o).reshape(h, w).astype(np.float32)
    o += w * h * 4
    cost = normals = None
    if flags & 1:
        cost = np.frombuffer(b, "<f4", w * h, o).reshape(h, w).astype(np.float32)
        o += w * h * 4
    if flags & 2:
        normals = np.frombuffer(b, "<f4", 3 * w * h, o).reshape(h, w, 3).astype(np.float32)
    return depth, cost, normals
