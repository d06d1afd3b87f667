"""Seeded synthetic scenes for the dense MVS path.

The reference ships no test data for its dense stage, so inputs are rendered
here: a corner room of mutually orthogonal planes (floor + walls), optional
seeded spheres and boxes, each textured with 4-octave value noise sampled
bilinearly (trilinearly for objects), rendered to uint8 grayscale by ray
casting through pinhole cameras (k = 0, cx = W/2, cy = H/2, as
``Cam::new`` in sfm-lite/src/geom.rs:23-25).  Poses are world->camera
(x_cam = R x_world + t), the convention of ``Pose`` (geom.rs:49-53).

Rendering runs in torch on any device (the GPU when present, which makes
the 1920x1080 scenes cheap); the images, not the renderer, are the inputs
that the HIP path and the oracle both consume, so renderer precision never
enters a parity comparison.  The exact ray-cast depth and normal maps are
kept as accuracy ground truth.

Scene construction (all units world units):
  * corner at the origin; floor y = 0 (normal +y), walls x = 0 and z = 0;
    walls are ``room`` units wide and ``wall_h`` tall; rays that miss every
    surface see a flat background (textureless, gated out by the census
    texture test);
  * cameras on a horizontal arc of radius ``radius`` around the vertical
    axis through ``arc_center`` (a point on the room diagonal), at height
    ``cam_h``, spread over ``arc_deg`` centred on the diagonal, all looking
    at ``target``;
  * optional flat-shaded rectangle on the x = 0 wall (``flat_patch``),
    and additive Gaussian noise ``noise_sigma`` (DN) before quantisation.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

__all__ = [
    "SceneSpec", "View", "Scene", "render_scene", "config_spec", "select_sources",
    "look_at",
]


@dataclass
class View:
    """One rendered view: intrinsics, world->camera pose, images, ground truth."""
    w: int
    h: int
    f: float
    cx: float
    cy: float
    k: float
    R: np.ndarray          # (3, 3) float64, world -> camera
    t: np.ndarray          # (3,) float64
    gray: np.ndarray       # (h, w) uint8
    rgb: Optional[np.ndarray] = None   # (h, w, 3) uint8 or None (= gray replicated)
    gt_depth: Optional[np.ndarray] = None   # (h, w) float32, 0 = background
    gt_normal: Optional[np.ndarray] = None  # (h, w, 3) float32, camera frame

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.t


@dataclass
class SceneSpec:
    seed: int = 1
    n_views: int = 10
    sizes: Sequence[Tuple[int, int]] = ((320, 240),)   # per view (w, h); cycled or sampled
    size_seeded: bool = False       # draw each view's size from `sizes` with the seed
    focal_rel: Optional[float] = None   # f = focal_rel * W when set
    focal: float = 275.0
    arc_deg: float = 60.0
    radius: float = 4.0
    arc_center: float = -1.0        # signed offset of the arc axis along the room diagonal
    cam_h: float = 1.6
    target: Tuple[float, float, float] = (0.5, 1.0, 0.5)
    room: float = 10.0
    wall_h: float = 6.0
    n_objects: int = 0
    obj_size: Tuple[float, float] = (0.2, 0.6)
    flat_patch: bool = False
    noise_sigma: float = 0.0
    tex_freq: float = 4.0           # lattice cells per world unit at the coarsest octave
    octaves: int = 4
    persistence: float = 0.65       # amplitude ratio between successive octaves
    obj_region: Tuple[float, float] = (0.3, 1.8)   # object centres: x, z range (corner side of the arc)
    background: int = 128
    room_kind: str = "corner"       # "corner": floor + walls x=0, z=0; "box": closed room |x|,|z| <= room/2, y in [0, wall_h]
    orbit: bool = False             # cameras on a full circle around the vertical axis through the origin
    flat_frac: float = 0.0          # fraction of objects rendered flat-shaded (textureless)


@dataclass
class Scene:
    spec: SceneSpec
    views: List[View]
    objects: list = field(default_factory=list)


def look_at(center: np.ndarray, target: np.ndarray, up=(0.0, 1.0, 0.0)) -> Tuple[np.ndarray, np.ndarray]:
    """World->camera (R, t) for a camera at `center` looking at `target`;
    camera axes x right, y down, z forward (COLMAP convention)."""
    fwd = np.asarray(target, np.float64) - np.asarray(center, np.float64)
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    R = np.stack([right, down, fwd])
    t = -R @ np.asarray(center, np.float64)
    return R, t


# --------------------------------------------------------------------------
# value noise
# --------------------------------------------------------------------------

_M32 = 0xFFFFFFFF


def _hash(ix, iy, iz, seed: int):
    h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761 & _M32)
    h = h & _M32
    h = (((h >> 16) ^ h) * 0x45D9F3B) & _M32
    h = (((h >> 16) ^ h) * 0x45D9F3B) & _M32
    h = (h >> 16) ^ h
    return (h & 0xFFFFFF).to(torch.float32) * (1.0 / 16777216.0)


def _noise2(s, t, seed: int, freq: float, octaves: int, persistence: float = 0.5):
    acc = torch.zeros_like(s, dtype=torch.float32)
    norm = 0.0
    amp = 1.0
    for o in range(octaves):
        fs = s * (freq * 2 ** o)
        ft = t * (freq * 2 ** o)
        i0 = torch.floor(fs)
        j0 = torch.floor(ft)
        ax = (fs - i0).to(torch.float32)
        ay = (ft - j0).to(torch.float32)
        i0 = i0.to(torch.int64)
        j0 = j0.to(torch.int64)
        z = torch.full_like(i0, o)
        h00 = _hash(i0, j0, z, seed)
        h10 = _hash(i0 + 1, j0, z, seed)
        h01 = _hash(i0, j0 + 1, z, seed)
        h11 = _hash(i0 + 1, j0 + 1, z, seed)
        v = (h00 * (1 - ax) + h10 * ax) * (1 - ay) + (h01 * (1 - ax) + h11 * ax) * ay
        acc += amp * v
        norm += amp
        amp *= persistence
    return acc / norm


def _noise3(p, seed: int, freq: float, octaves: int, persistence: float = 0.5):
    acc = torch.zeros(p.shape[:-1], dtype=torch.float32, device=p.device)
    norm = 0.0
    amp = 1.0
    for o in range(octaves):
        q = p * (freq * 2 ** o)
        q0 = torch.floor(q)
        a = (q - q0).to(torch.float32)
        q0 = q0.to(torch.int64)
        v = torch.zeros_like(acc)
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    wgt = ((a[..., 0] if cx else 1 - a[..., 0]) * (a[..., 1] if cy else 1 - a[..., 1])
                           * (a[..., 2] if cz else 1 - a[..., 2]))
                    v += wgt * _hash(q0[..., 0] + cx, q0[..., 1] + cy, q0[..., 2] + cz + 7919 * o, seed)
        acc += amp * v
        norm += amp
        amp *= persistence
    return acc / norm


def _tone(n, persistence=0.5):
    # value-noise sums concentrate around 0.5; stretch to use the u8 range
    gain = 330.0 if persistence <= 0.5 else 420.0
    return torch.clamp(128.0 + gain * (n - 0.5), 0.0, 255.0)


# --------------------------------------------------------------------------
# scene construction
# --------------------------------------------------------------------------

def config_spec(name: str) -> SceneSpec:
    """Scene specs of the named workloads (seeds recorded here)."""
    if name == "corner320":      # 10 x 320x240, f=275, 60 deg arc, radius 4
        return SceneSpec(seed=1, n_views=10, sizes=((320, 240),), focal=275.0, arc_deg=60.0,
                         radius=4.0, arc_center=-0.5, tex_freq=8.0)
    if name == "room1080":       # 20 x 1920x1080, f=1650, 90 deg arc, 12 objects
        return SceneSpec(seed=2, n_views=20, sizes=((1920, 1080),), focal=1650.0, arc_deg=90.0,
                         radius=4.0, arc_center=0.85, n_objects=12, flat_patch=True, noise_sigma=2.0,
                         tex_freq=24.0, persistence=0.65)
    if name == "mixed64":        # 64 views of uneven size, 30 objects
        return SceneSpec(seed=3, n_views=64, sizes=((640, 480), (1280, 720), (1920, 1080), (2560, 1440)),
                         size_seeded=True, focal_rel=0.86, arc_deg=90.0, radius=4.0, arc_center=0.85,
                         n_objects=30, flat_patch=True, noise_sigma=2.0, tex_freq=24.0, persistence=0.65)
    if name == "orbit100":       # 100 x 1920x1080 on a 360 deg orbit, radius 5, closed 6-plane room, 50 objects
        return SceneSpec(seed=5, n_views=100, sizes=((1920, 1080),), focal=1650.0, radius=5.0, orbit=True,
                         cam_h=1.6, target=(0.0, 0.9, 0.0), room_kind="box", room=16.0, wall_h=5.0, n_objects=50,
                         obj_region=(-2.5, 2.5), flat_frac=0.1, noise_sigma=2.0, tex_freq=24.0, persistence=0.65)
    if name == "tiny":           # small scene for fast parity tests
        return SceneSpec(seed=11, n_views=5, sizes=((64, 48),), focal=60.0, arc_deg=40.0,
                         radius=4.0, arc_center=-1.0, n_objects=2, tex_freq=3.0)
    raise KeyError(name)


def _objects(spec: SceneSpec, rng: np.random.Generator):
    objs = []
    for i in range(spec.n_objects):
        kind = "sphere" if rng.random() < 0.5 else "box"
        size = float(rng.uniform(*spec.obj_size))
        # placed in the corner region, >= ~2 units from every camera, resting on or above the floor
        x = float(rng.uniform(*spec.obj_region))
        z = float(rng.uniform(*spec.obj_region))
        y = size + float(rng.uniform(0.0, 1.0))
        flat = bool(rng.random() < spec.flat_frac)
        objs.append({"kind": kind, "c": (x, y, z), "s": size, "seed": int(spec.seed * 1000 + 500 + i), "flat": flat})
    return objs


def _room_planes(spec: SceneSpec):
    """(axis, coordinate, normal sign) of the room's planes."""
    if spec.room_kind == "box":
        L = spec.room / 2
        return [(1, 0.0, 1.0), (0, -L, 1.0), (2, -L, 1.0), (0, L, -1.0), (2, L, -1.0), (1, spec.wall_h, -1.0)]
    return [(1, 0.0, 1.0), (0, 0.0, 1.0), (2, 0.0, 1.0)]


def _room_extent(spec: SceneSpec, axis: int):
    if spec.room_kind == "box":
        return (0.0, spec.wall_h) if axis == 1 else (-spec.room / 2, spec.room / 2)
    return (0.0, spec.wall_h) if axis == 1 else (0.0, spec.room)


def _view_sizes(spec: SceneSpec, rng: np.random.Generator):
    if spec.size_seeded:
        idx = rng.integers(0, len(spec.sizes), size=spec.n_views)
        return [tuple(spec.sizes[i]) for i in idx]
    return [tuple(spec.sizes[i % len(spec.sizes)]) for i in range(spec.n_views)]


def render_scene(spec: SceneSpec, device: Optional[str] = None, with_gt: bool = True) -> Scene:
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    rng = np.random.default_rng(spec.seed)
    sizes = _view_sizes(spec, rng)
    objs = _objects(spec, rng)
    diag = np.array([1.0, 0.0, 1.0]) / math.sqrt(2.0)
    axis = diag * spec.arc_center
    views = []
    for vi in range(spec.n_views):
        w, h = sizes[vi]
        f = spec.focal_rel * w if spec.focal_rel else spec.focal
        if spec.orbit:
            th = 2.0 * math.pi * vi / spec.n_views
            c = spec.radius * np.array([math.cos(th), 0.0, math.sin(th)])
            c[1] = spec.cam_h
            R, t = look_at(c, np.array(spec.target))
        else:
            frac = 0.5 if spec.n_views == 1 else vi / (spec.n_views - 1)
            th = math.radians(45.0 - spec.arc_deg / 2 + spec.arc_deg * frac)
            c = axis + spec.radius * np.array([math.cos(th), 0.0, math.sin(th)])
            c[1] = spec.cam_h
            R, t = look_at(c, np.array(spec.target))
        gray, depth, normal = _render_view(spec, objs, w, h, f, R, t, device,
                                           noise_seed=spec.seed * 1000 + vi)
        v = View(w=w, h=h, f=float(f), cx=w * 0.5, cy=h * 0.5, k=0.0, R=R, t=t, gray=gray)
        if with_gt:
            v.gt_depth = depth
            v.gt_normal = normal
        views.append(v)
    return Scene(spec=spec, views=views, objects=objs)


def _render_view(spec, objs, w, h, f, R, t, device, noise_seed):
    dt = torch.float64 if device == "cpu" else torch.float32
    Rt = torch.tensor(R, dtype=dt, device=device)
    C = torch.tensor(-R.T @ t, dtype=dt, device=device)
    us = torch.arange(w, dtype=dt, device=device)
    vs = torch.arange(h, dtype=dt, device=device)
    vv, uu = torch.meshgrid(vs, us, indexing="ij")
    dcam = torch.stack([(uu - w * 0.5) / f, (vv - h * 0.5) / f, torch.ones_like(uu)], -1)
    dw = dcam @ Rt            # row vector times R == R^T d
    INF = float("inf")
    best = torch.full((h, w), INF, dtype=dt, device=device)
    nrm = torch.zeros((h, w, 3), dtype=dt, device=device)
    sid = torch.full((h, w), -1, dtype=torch.int64, device=device)   # surface id

    def take(tt, n, idv):
        nonlocal best, nrm, sid
        m = (tt > 1e-4) & (tt < best)
        best = torch.where(m, tt, best)
        nrm = torch.where(m[..., None], n, nrm)
        sid = torch.where(m, torch.full_like(sid, idv), sid)

    # room planes (axis, coordinate, normal sign); ids 0-2 are the corner planes
    for pid, (a, coord, sgn) in enumerate(_room_planes(spec)):
        da = dw[..., a]
        with torch.no_grad():
            tt = torch.where(da.abs() > 1e-12, (coord - C[a]) / da, torch.full_like(da, INF))
        p = C + dw * tt[..., None]
        inside = torch.ones_like(tt, dtype=torch.bool)
        for b in range(3):
            if b == a:
                continue
            lo, hi = _room_extent(spec, b)
            inside &= (p[..., b] >= lo - 1e-9) & (p[..., b] <= hi + 1e-9)
        tt = torch.where(inside, tt, torch.full_like(tt, INF))
        n = torch.zeros(3, dtype=dt, device=device)
        n[a] = sgn
        take(tt, n.expand(h, w, 3), pid)
    for oi, o in enumerate(objs):
        c = torch.tensor(o["c"], dtype=dt, device=device)
        s = o["s"]
        if o["kind"] == "sphere":
            oc = C - c
            b = (dw * oc).sum(-1)
            a2 = (dw * dw).sum(-1)
            cc = (oc * oc).sum() - s * s
            disc = b * b - a2 * cc
            sq = torch.sqrt(torch.clamp(disc, min=0.0))
            tt = (-b - sq) / a2
            tt = torch.where(disc > 0, tt, torch.full_like(tt, INF))
            p = C + dw * tt[..., None]
            n = (p - c) / s
        else:
            inv = 1.0 / torch.where(dw.abs() > 1e-12, dw, torch.full_like(dw, 1e-12))
            t0 = (c - s - C) * inv
            t1 = (c + s - C) * inv
            tmin = torch.minimum(t0, t1)
            tmax = torch.maximum(t0, t1)
            tn, ax = tmin.max(-1)
            tf = tmax.min(-1).values
            tt = torch.where((tn <= tf) & (tn > 0), tn, torch.full_like(tn, INF))
            n = torch.nn.functional.one_hot(ax, 3).to(dt) * -torch.sign(torch.gather(dw, -1, ax[..., None]))
        take(tt, n, len(_room_planes(spec)) + oi)
    hit = torch.isfinite(best)
    p = C + dw * torch.where(hit, best, torch.zeros_like(best))[..., None]
    val = torch.full((h, w), float(spec.background), dtype=torch.float32, device=device)
    # planes: 2D value noise in the plane's own coordinates
    for pid, (a, _, _) in enumerate(_room_planes(spec)):
        b1, b2 = [b for b in (0, 1, 2) if b != a][:2] if a != 1 else (0, 2)
        m = sid == pid
        if m.any():
            nv = _noise2(p[..., b1][m], p[..., b2][m], spec.seed * 1000 + 100 + pid, spec.tex_freq, spec.octaves,
                         spec.persistence)
            val[m] = _tone(nv, spec.persistence)
    nplanes = len(_room_planes(spec))
    for oi, o in enumerate(objs):
        m = sid == nplanes + oi
        if m.any() and o.get("flat"):
            val[m] = 90.0 + 60.0 * (oi % 3)
            continue
        if m.any():
            q = p[m] - torch.tensor(o["c"], dtype=dt, device=device)
            nv = _noise3(q + 50.0, o["seed"], spec.tex_freq, spec.octaves, spec.persistence)
            val[m] = _tone(nv, spec.persistence)
    if spec.flat_patch:
        # textureless rectangle on the x = 0 wall
        m = (sid == 1) & (p[..., 2] >= 1.6) & (p[..., 2] <= 2.4) & (p[..., 1] >= 1.2) & (p[..., 1] <= 2.0)
        val[m] = 150.0
    if spec.noise_sigma > 0:
        g = torch.Generator(device=device)
        g.manual_seed(int(noise_seed))
        val = val + spec.noise_sigma * torch.randn((h, w), generator=g, device=device, dtype=torch.float32)
    gray = torch.clamp(torch.round(val), 0, 255).to(torch.uint8)
    depth = torch.where(hit, best, torch.zeros_like(best)).to(torch.float32)   # dcam.z == 1
    ncam = (nrm @ Rt.T).to(torch.float32)
    return gray.cpu().numpy(), depth.cpu().numpy(), ncam.cpu().numpy()


def select_sources(views: Sequence[View], k: int) -> List[List[int]]:
    """Per reference view, the k other views with the smallest angle between
    viewing directions (ties by index).  The reference picks sources by
    shared sparse points (mvs/src/lib.rs:326-342); explicit lists keep the
    harness independent of that tie order."""
    dirs = [v.R[2] for v in views]
    out = []
    for i in range(len(views)):
        cand = []
        for j in range(len(views)):
            if j == i:
                continue
            ang = math.acos(max(-1.0, min(1.0, float(np.dot(dirs[i], dirs[j])))))
            cand.append((ang, j))
        cand.sort()
        out.append([j for _, j in cand[:k]])
    return out


def uneven_jobs(sc: Scene, zmin: float = 1.0, zmax: float = 10.0):
    """Stereo jobs of the uneven batch: per view a support-set size K drawn
    uniformly from [4, 12] with the scene's seed + 100, its K nearest views
    by viewing direction as sources.  Returns [(ref, srcs, zmin, zmax)]."""
    rng = np.random.default_rng(sc.spec.seed + 100)
    ks = rng.integers(4, 13, size=len(sc.views))
    return [(r, select_sources(sc.views, int(ks[r]))[r], zmin, zmax) for r in range(len(sc.views))]


def sparse_points(sc: Scene, n: int, seed: int = 0, rel_tol: float = 0.01):
    """A synthetic sparse reconstruction: n surface points sampled from the
    ground-truth depth of random views, each with the track of views that
    see it (projection inside the image and ground-truth depth within
    rel_tol), as COLMAP's points3D would carry.  Returns (xyz (n, 3), tracks)."""
    rng = np.random.default_rng(seed)
    views = sc.views
    pts, tracks = [], []
    tries = 0
    while len(pts) < n and tries < 50 * n:
        tries += 1
        vi = int(rng.integers(len(views)))
        v = views[vi]
        x, y = int(rng.integers(v.w)), int(rng.integers(v.h))
        z = float(v.gt_depth[y, x])
        if z <= 0:
            continue
        Xc = np.array([(x - v.cx) / v.f * z, (y - v.cy) / v.f * z, z])
        Xw = v.R.T @ (Xc - v.t)
        tr = []
        for ui, u in enumerate(views):
            q = u.R @ Xw + u.t
            if q[2] <= 1e-6:
                continue
            px, py = u.f * q[0] / q[2] + u.cx, u.f * q[1] / q[2] + u.cy
            ix, iy = int(round(px)), int(round(py))
            if 0 <= ix < u.w and 0 <= iy < u.h:
                zu = float(u.gt_depth[iy, ix])
                if zu > 0 and abs(zu - q[2]) <= rel_tol * q[2]:
                    tr.append(ui)
        if len(tr) >= 2:
            pts.append(Xw)
            tracks.append(tr)
    return np.asarray(pts, np.float64).reshape(-1, 3), tracks
