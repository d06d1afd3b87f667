"""Host-side mirror of the reference `mvs` crate interface over the HIP C ABI.

Names, argument meaning and defaults follow the reference:
  MvsOptions        mvs/src/lib.rs:33-70
  ViewData          mvs/src/lib.rs:86-101   (gray kept as u8; f32 gray = u8/255)
  census_map        mvs/src/lib.rs:119-142
  DepthMap          mvs/src/stereo.rs:33-38 (plus the accepted normals)
  estimate_depth    mvs/src/stereo.rs:234-241
  fuse_depth_maps   mvs/src/fuse.rs:64-71
  write_ply_binary  mvs/src/fuse.rs:221-239
  Cam / Pose        sfm-lite/src/geom.rs:13-76

Every compute call runs the HIP kernels of libmvs_hip.so; there is no CPU
fallback.  One `MvsContext` owns one device stream and the device-resident
views; module-level functions use a lazily created context on device 0.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import MvsJobC, MvsOptionsC, MvsPointC, MvsStatsC, MvsViewC, check, ptr

POINT_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("rgb", "u1", (3,)), ("pad", "u1", (5,))])


@dataclass
class MvsOptions:
    """MvsOptions with the reference defaults (lib.rs:56-70)."""
    max_dim: int = 900
    num_sources: int = 3
    sweeps: int = 4
    num_hyp: int = 48
    min_consistent: int = 2
    depth_diff: float = 0.03
    cost_thr: float = 0.38
    voxel_rel: float = 0.0015
    save_depth_vis: bool = False

    def to_c(self) -> MvsOptionsC:
        o = MvsOptionsC()
        o.max_dim = int(self.max_dim)
        o.num_sources = int(self.num_sources)
        o.sweeps = int(self.sweeps)
        o.num_hyp = int(self.num_hyp)
        o.min_consistent = int(self.min_consistent)
        o.save_depth_vis = int(bool(self.save_depth_vis))
        o.depth_diff = float(self.depth_diff)
        o.voxel_rel = float(self.voxel_rel)
        o.cost_thr = float(self.cost_thr)
        return o


@dataclass
class Cam:
    """SIMPLE_RADIAL camera (geom.rs:13-21)."""
    f: float
    cx: float
    cy: float
    k: float = 0.0


@dataclass
class Pose:
    """World->camera pose x_cam = R x + t (geom.rs:49-53)."""
    r: np.ndarray
    t: np.ndarray


@dataclass
class ViewData:
    """A registered view at working resolution (lib.rs:86-101)."""
    w: int
    h: int
    cam: Cam
    pose: Pose
    gray: np.ndarray                 # (h, w) uint8
    rgb: Optional[np.ndarray] = None  # (h, w, 3) uint8 or None
    name: str = ""
    image_id: int = 0

    @classmethod
    def from_scene_view(cls, v, name: str = "", image_id: int = 0) -> "ViewData":
        return cls(w=v.w, h=v.h, cam=Cam(v.f, v.cx, v.cy, v.k), pose=Pose(np.asarray(v.R, np.float64),
                                                                           np.asarray(v.t, np.float64)),
                   gray=np.ascontiguousarray(v.gray, np.uint8),
                   rgb=None if v.rgb is None else np.ascontiguousarray(v.rgb, np.uint8),
                   name=name, image_id=image_id)


@dataclass
class DepthMap:
    """depth (0 = invalid) and cost per pixel (stereo.rs:33-38); `normals`
    carries the accepted unit normals (camera frame) as an extra."""
    depth: np.ndarray
    cost: np.ndarray
    normals: Optional[np.ndarray] = None


@dataclass
class MvsStats:
    n_eval_init: int = 0
    n_eval_pass: int = 0
    n_pixels: int = 0
    taps_per_eval_sum: int = 0
    ms_prepare: float = 0.0
    ms_init: float = 0.0
    ms_pass: float = 0.0
    ms_gate: float = 0.0
    ms_fuse: float = 0.0
    n_candidates: int = 0
    n_points: int = 0
    n_pass_launches: int = 0
    graph_used: int = 0
    ms_learned_volume: float = 0.0
    ms_learned_net: float = 0.0
    tile_w: int = 0
    tile_h: int = 0
    persistent: int = 0
    evaluator: int = 0


def views_c(views: Sequence[ViewData]):
    """mvs_view_t array (and the arrays it points into) for `views`."""
    arr = (MvsViewC * len(views))()
    keep = []
    for i, v in enumerate(views):
        g = np.ascontiguousarray(v.gray, np.uint8)
        if g.shape != (v.h, v.w):
            raise ValueError("view %d: gray shape %s != (h, w) = (%d, %d)" % (i, g.shape, v.h, v.w))
        keep.append(g)
        c = arr[i]
        c.w, c.h = int(v.w), int(v.h)
        c.f, c.cx, c.cy, c.k = float(v.cam.f), float(v.cam.cx), float(v.cam.cy), float(v.cam.k)
        R = np.asarray(v.pose.r, np.float64).reshape(9)
        t = np.asarray(v.pose.t, np.float64).reshape(3)
        for j in range(9):
            c.R[j] = float(R[j])
        for j in range(3):
            c.t[j] = float(t[j])
        c.gray = g.ctypes.data_as(C.POINTER(C.c_uint8))
        name = v.name.encode("utf-8") if v.name else None   # None: the library names it view<i>
        keep.append(name)
        c.name = name
        if v.rgb is not None:
            rgb = np.ascontiguousarray(v.rgb, np.uint8)
            if rgb.size != 3 * v.w * v.h:
                raise ValueError("view %d: rgb has %d bytes, expected %d" % (i, rgb.size, 3 * v.w * v.h))
            keep.append(rgb)
            c.rgb = rgb.ctypes.data_as(C.POINTER(C.c_uint8))
        else:
            c.rgb = None
    return arr, keep



class MvsContext:
    """One device, one HIP stream, the device-resident views."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.mvs_ctx_create(int(device), C.byref(h)), "mvs_ctx_create")
        self._h = h
        self.views: Optional[List[ViewData]] = None
        self._keep = []

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mvs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- log -------------------------------------------------------------------
    def set_log(self, log: Optional[Callable[[str], None]]):
        """The reference's `log` callback (lib.rs:299, fuse.rs:70): receives
        the progress lines of run_mvs and fusion; None switches it off.
        Returns the previous callback."""
        prev = getattr(self, "_log_py", None)
        if log is None:
            cb = _lib.LOG_FN()
        else:
            cb = _lib.LOG_FN(lambda line, _user: log(line.decode("utf-8", "replace")))
        check(self._lib.mvs_set_log(self._h, cb, None), "mvs_set_log")
        self._log_cb = cb          # the C side keeps only the pointer
        self._log_py = log
        return prev

    # -- views ---------------------------------------------------------------
    def upload_views(self, views: Sequence[ViewData]):
        arr, keep = views_c(views)
        check(self._lib.mvs_upload_views(self._h, arr, len(views)), "mvs_upload_views")
        self.views = list(views)

    def ensure_views(self, views: Sequence[ViewData]):
        if self.views is None or len(self.views) != len(views) or any(a is not b for a, b in zip(self.views, views)):
            self.upload_views(views)

    def view_census(self, i: int) -> np.ndarray:
        v = self.views[i]
        out = np.empty((v.h, v.w), np.uint64)
        check(self._lib.mvs_view_census(self._h, int(i), ptr(out)), "mvs_view_census")
        return out

    # -- census ----------------------------------------------------------------
    def census(self, gray: np.ndarray) -> np.ndarray:
        g = np.ascontiguousarray(gray, np.uint8)
        h, w = g.shape
        out = np.empty((h, w), np.uint64)
        check(self._lib.mvs_census(self._h, ptr(g), int(w), int(h), ptr(out)), "mvs_census")
        return out

    # -- stereo ----------------------------------------------------------------
    def estimate_depth(self, ref_idx: int, srcs: Sequence[int], zmin: float, zmax: float, opts: MvsOptions,
                       max_passes: int = -1, want_planes: bool = False, want_normals: bool = True):
        v = self.views[ref_idx]
        s = np.ascontiguousarray(srcs, np.int32)
        depth = np.empty((v.h, v.w), np.float32)
        cost = np.empty((v.h, v.w), np.float32)
        normals = np.empty((v.h, v.w, 3), np.float32) if want_normals else None
        planes = np.empty((v.h, v.w, 5), np.float32) if want_planes else None
        n_passes = 2 * max(int(opts.sweeps), 1)
        n_eval = np.zeros(1 + n_passes, np.int64)
        oc = opts.to_c()
        check(self._lib.mvs_estimate_depth_ex(self._h, int(ref_idx), ptr(s), len(s), float(zmin), float(zmax),
                                              C.byref(oc), int(max_passes), ptr(depth), ptr(cost), ptr(normals),
                                              ptr(planes), ptr(n_eval)), "mvs_estimate_depth")
        dm = DepthMap(depth=depth, cost=cost, normals=normals)
        if want_planes:
            return dm, planes, n_eval
        return dm

    def estimate_depth_batch(self, jobs, opts: MvsOptions, mode: int = 0, download: bool = True):
        """jobs: sequence of (ref, srcs, zmin, zmax).  Returns {ref: DepthMap}
        when `download`, else None (results stay on the device)."""
        arr = (MvsJobC * len(jobs))()
        keep = []
        out = {}
        for i, (ref, srcs, zmin, zmax) in enumerate(jobs):
            s = np.ascontiguousarray(srcs, np.int32)
            keep.append(s)
            j = arr[i]
            j.ref, j.k = int(ref), len(s)
            j.srcs = s.ctypes.data_as(C.POINTER(C.c_int32))
            j.zmin, j.zmax = float(zmin), float(zmax)
            if download:
                v = self.views[ref]
                d = np.empty((v.h, v.w), np.float32)
                c = np.empty((v.h, v.w), np.float32)
                n = np.empty((v.h, v.w, 3), np.float32)
                keep += [d, c, n]
                j.depth = d.ctypes.data_as(C.POINTER(C.c_float))
                j.cost = c.ctypes.data_as(C.POINTER(C.c_float))
                j.normals = n.ctypes.data_as(C.POINTER(C.c_float))
                out[int(ref)] = DepthMap(depth=d, cost=c, normals=n)
            else:
                j.depth = j.cost = j.normals = None
        oc = opts.to_c()
        check(self._lib.mvs_estimate_depth_batch(self._h, arr, len(jobs), C.byref(oc), int(mode)),
              "mvs_estimate_depth_batch")
        return out if download else None

    def autotune(self, jobs, opts: MvsOptions, passes: int = 2):
        """Time the launch-shape candidates on `jobs` and keep the fastest
        for their shape class.  Returns (tile_w, tile_h, persistent, evaluator)."""
        arr = (MvsJobC * len(jobs))()
        keep = []
        for i, (ref, srcs, zmin, zmax) in enumerate(jobs):
            s = np.ascontiguousarray(srcs, np.int32)
            keep.append(s)
            arr[i].ref, arr[i].k = int(ref), len(s)
            arr[i].srcs = s.ctypes.data_as(C.POINTER(C.c_int32))
            arr[i].zmin, arr[i].zmax = float(zmin), float(zmax)
        best = np.zeros(4, np.int32)
        ms = np.zeros(32, np.float32)
        cfg = np.zeros((32, 4), np.int32)
        nc = np.zeros(1, np.int32)
        oc = opts.to_c()
        check(self._lib.mvs_autotune(self._h, arr, len(jobs), C.byref(oc), int(passes), ptr(best), ptr(ms), ptr(cfg),
                                     ptr(nc)), "mvs_autotune")
        self.last_tune = [("%dx%d%s%s" % (c[0], c[1], ["", "p", "s"][c[2]], "r.vwx"[c[3]] if c[3] != 1 else ""),
                           float(t)) for c, t in zip(cfg[:nc[0]], ms[:nc[0]])]
        return tuple(int(b) for b in best)

    def get_depth(self, view: int) -> DepthMap:
        v = self.views[view]
        d = np.empty((v.h, v.w), np.float32)
        c = np.empty((v.h, v.w), np.float32)
        n = np.empty((v.h, v.w, 3), np.float32)
        check(self._lib.mvs_get_depth(self._h, int(view), ptr(d), ptr(c), ptr(n)), "mvs_get_depth")
        return DepthMap(d, c, n)

    # -- fusion ----------------------------------------------------------------
    @staticmethod
    def _orders(nv: int, sources: Sequence[Sequence[int]]):
        off = np.zeros(nv + 1, np.int32)
        flat = []
        for i in range(nv):
            o = list(sources[i]) if i < len(sources) else []
            flat += o
            off[i + 1] = off[i] + len(o)
        return off, np.asarray(flat if flat else [0], np.int32)

    def _points(self, pp, n) -> np.ndarray:
        n = int(n.value)
        out = np.zeros(n, POINT_DTYPE)
        if n:
            C.memmove(out.ctypes.data, pp, n * C.sizeof(MvsPointC))
        self._lib.mvs_free_points(pp)
        return out

    def fuse_resident(self, has_depth: Sequence[bool], sources: Sequence[Sequence[int]], cell: float,
                      opts: MvsOptions) -> np.ndarray:
        nv = len(self.views)
        hd = np.asarray([1 if h else 0 for h in has_depth], np.int32)
        off, flat = self._orders(nv, sources)
        pp = C.POINTER(MvsPointC)()
        n = C.c_int64()
        oc = opts.to_c()
        check(self._lib.mvs_fuse(self._h, ptr(hd), ptr(off), ptr(flat), float(cell), C.byref(oc), C.byref(pp),
                                 C.byref(n)), "mvs_fuse")
        return self._points(pp, n)

    def fuse_depth_maps(self, depth_maps: Sequence[Optional[DepthMap]], sources: Sequence[Sequence[int]],
                        cell: float, opts: MvsOptions) -> np.ndarray:
        nv = len(self.views)
        arr = (C.c_void_p * nv)()
        keep = []
        for i in range(nv):
            dm = depth_maps[i]
            if dm is None:
                arr[i] = None
            else:
                d = np.ascontiguousarray(dm.depth, np.float32)
                keep.append(d)
                arr[i] = d.ctypes.data
        off, flat = self._orders(nv, sources)
        pp = C.POINTER(MvsPointC)()
        n = C.c_int64()
        oc = opts.to_c()
        check(self._lib.mvs_fuse_depth_maps(self._h, arr, ptr(off), ptr(flat), float(cell), C.byref(oc),
                                            C.byref(pp), C.byref(n)), "mvs_fuse_depth_maps")
        return self._points(pp, n)

    def stats(self) -> MvsStats:
        s = MvsStatsC()
        check(self._lib.mvs_get_stats(self._h, C.byref(s)), "mvs_get_stats")
        return MvsStats(**{k: getattr(s, k) for k, _ in MvsStatsC._fields_})

    # -- instrumentation -------------------------------------------------------
    def debug_plane_cost(self, ref: int, srcs: Sequence[int], queries: np.ndarray):
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, 7)
        n = len(q)
        s = np.ascontiguousarray(srcs, np.int32)
        cost = np.empty(n, np.float32)
        ham = np.empty(n, np.int32)
        ninv = np.empty(n, np.int32)
        peak = np.empty(n, np.float32)
        check(self._lib.mvs_debug_plane_cost(self._h, int(ref), ptr(s), len(s), ptr(q), n, ptr(cost), ptr(ham),
                                             ptr(ninv), ptr(peak)), "mvs_debug_plane_cost")
        return cost, ham, ninv, peak

    def debug_run_passes(self, ref: int, srcs: Sequence[int], zmin: float, zmax: float, planes: np.ndarray,
                         pass_lo: int, pass_hi: int, n_passes: int):
        """Passes pass_lo..pass_hi-1 (of n_passes) from the given (h, w, 5)
        planes instead of the init sweep; returns (planes, n_eval per pass)."""
        v = self.views[ref]
        s = np.ascontiguousarray(srcs, np.int32)
        pl = np.ascontiguousarray(planes, np.float32).reshape(v.h, v.w, 5).copy()
        ne = np.zeros(max(pass_hi - pass_lo, 1), np.int64)
        check(self._lib.mvs_debug_run_passes(self._h, int(ref), ptr(s), len(s), float(zmin), float(zmax),
                                             int(pass_lo), int(pass_hi), int(n_passes), ptr(pl), ptr(ne)),
              "mvs_debug_run_passes")
        return pl, ne[: pass_hi - pass_lo]

    def debug_rng(self, xy: np.ndarray, salt: np.ndarray, ndraw: int):
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        salt = np.ascontiguousarray(salt, np.uint64).reshape(-1)
        n = len(xy)
        seed = np.empty(n, np.uint64)
        uni = np.empty((n, ndraw), np.float32)
        check(self._lib.mvs_debug_rng(self._h, ptr(xy), ptr(salt), n, int(ndraw), ptr(seed), ptr(uni)),
              "mvs_debug_rng")
        return seed, uni


_default_ctx: Optional[MvsContext] = None
_ctx_lock = threading.Lock()


def default_context() -> MvsContext:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = MvsContext(0)
        return _default_ctx


# ---------------------------------------------------------------------------
# reference-shaped free functions
# ---------------------------------------------------------------------------

def census_map(gray: np.ndarray, w: int, h: int, ctx: Optional[MvsContext] = None) -> np.ndarray:
    """census_map (lib.rs:119-142) of a u8 image (h*w), on the GPU; returns u64[h*w]."""
    g = np.asarray(gray)
    if g.dtype != np.uint8:
        # ViewData.gray is u8/255 in f32; recover the u8 exactly
        g = np.rint(np.asarray(g, np.float64) * 255.0).astype(np.uint8)
    g = g.reshape(h, w)
    return (ctx or default_context()).census(g).reshape(-1)


def estimate_depth(views: Sequence[ViewData], ref_idx: int, srcs: Sequence[int], zmin: float, zmax: float,
                   opts: MvsOptions, ctx: Optional[MvsContext] = None) -> DepthMap:
    """stereo::estimate_depth (stereo.rs:234-394) on the GPU."""
    ctx = ctx or default_context()
    ctx.ensure_views(views)
    return ctx.estimate_depth(ref_idx, srcs, zmin, zmax, opts)


def fusion_cell(sparse: np.ndarray, voxel_rel: float) -> float:
    """Robust voxel cell of fuse_depth_maps (fuse.rs:72-88)."""
    lib = _lib.load()
    xyz = np.ascontiguousarray(sparse, np.float64).reshape(-1, 3)
    if len(xyz) == 0:
        raise ValueError("fusion_cell: no sparse points (the reference panics here)")
    return float(lib.mvs_fusion_cell(ptr(xyz), len(xyz), float(voxel_rel)))


def fuse_depth_maps(views: Sequence[ViewData], depth_maps: Sequence[Optional[DepthMap]],
                    sources: Sequence[Sequence[int]], sparse: Optional[np.ndarray], opts: MvsOptions,
                    log: Optional[Callable[[str], None]] = None, ctx: Optional[MvsContext] = None,
                    cell: Optional[float] = None) -> np.ndarray:
    """fuse::fuse_depth_maps (fuse.rs:64-215) on the GPU.  Returns a
    structured array (x, y, z, rgb) in first-insertion voxel order.  `cell`
    overrides the robust sparse-diagonal cell (absolute voxel size)."""
    ctx = ctx or default_context()
    ctx.ensure_views(views)
    if cell is None:
        cell = fusion_cell(sparse, opts.voxel_rel)
        if log:
            xyz = np.ascontiguousarray(sparse, np.float64).reshape(-1, 3)
            diag = float(_lib.load().mvs_fusion_diag(ptr(xyz), len(xyz)))
            log("mvs: fusion voxel cell %.4f (robust sparse diag %.2f)" % (cell, diag))
    prev = ctx.set_log(log) if log else None
    try:
        pts = ctx.fuse_depth_maps(depth_maps, sources, cell, opts)   # logs the candidate count
    finally:
        if log:
            ctx.set_log(prev)
    return pts


def write_ply_binary(path: str, pts: np.ndarray):
    """write_ply_binary (fuse.rs:221-239)."""
    lib = _lib.load()
    p = np.ascontiguousarray(pts, POINT_DTYPE)
    check(lib.mvs_write_ply(str(path).encode(), ptr(p), len(p)), "mvs_write_ply")


def _jobs_c(jobs):
    arr = (MvsJobC * len(jobs))()
    keep = []
    for i, jb in enumerate(jobs):
        ref, srcs, zmin, zmax = jb[:4]
        s = np.ascontiguousarray(srcs, np.int32)
        keep.append(s)
        arr[i].ref, arr[i].k = int(ref), len(s)
        arr[i].srcs = s.ctypes.data_as(C.POINTER(C.c_int32))
        arr[i].zmin, arr[i].zmax = float(zmin), float(zmax)
    return arr, keep


NET_STRIDE = 2   # MVS_NET_STRIDE: the network's grid is every second pixel


def coarse_shape(h: int, w: int, stride: int = NET_STRIDE):
    return (h + stride - 1) // stride, (w + stride - 1) // stride


class ProposalNet:
    """The learned plane-proposal network (BF16 U-Net on MFMA) of a context.
    It works on the coarse grid of every NET_STRIDE-th pixel: targets and
    predictions are (ceil(h/2), ceil(w/2), 5)."""

    def __init__(self, ctx: MvsContext, seed: int = 0):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        check(self._lib.mvs_net_create(ctx._h, C.c_uint64(seed), C.byref(h)), "mvs_net_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mvs_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(self._lib.mvs_net_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        check(self._lib.mvs_net_get_params(self._h, ptr(out)), "mvs_net_get_params")
        return out

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.num_params:
            raise ValueError("expected %d parameters, got %d" % (self.num_params, p.size))
        check(self._lib.mvs_net_set_params(self._h, ptr(p)), "mvs_net_set_params")

    def add_samples(self, jobs, targets) -> int:
        arr, keep = _jobs_c(jobs)
        tg = [np.ascontiguousarray(t, np.float32) for t in targets]
        tp = (C.c_void_p * len(tg))(*[t.ctypes.data for t in tg])
        r = self._lib.mvs_net_add_samples(self._h, arr, len(jobs), tp)
        check(r, "mvs_net_add_samples")
        return int(r)

    def clear_samples(self):
        check(self._lib.mvs_net_clear_samples(self._h), "mvs_net_clear_samples")

    def train_step(self, idx, lr: float):
        i = np.ascontiguousarray(idx, np.int32)
        loss = np.zeros(4, np.float32)
        check(self._lib.mvs_net_train_step(self._h, ptr(i), len(i), C.c_float(lr), ptr(loss)), "mvs_net_train_step")
        return loss

    def predict(self, job) -> np.ndarray:
        arr, keep = _jobs_c([job])
        v = self.ctx.views[job[0]]
        out = np.empty(coarse_shape(v.h, v.w) + (5,), np.float32)
        check(self._lib.mvs_net_predict(self._h, arr, ptr(out)), "mvs_net_predict")
        return out

    def debug_grads(self, inp: np.ndarray, target: np.ndarray, h: int, w: int):
        x = np.ascontiguousarray(inp, np.float32)
        n, Hp, Wp, c = x.shape
        assert c == 64
        t = np.ascontiguousarray(target, np.float32)
        grads = np.empty(self.num_params, np.float32)
        pred = np.empty((n, h, w, 5), np.float32)
        loss = np.empty(4, np.float32)
        check(self._lib.mvs_net_debug_grads(self._h, ptr(x), ptr(t), n, Hp, Wp, h, w, ptr(grads), ptr(pred),
                                            ptr(loss)), "mvs_net_debug_grads")
        return grads, pred, loss


def resize_image(ctx: MvsContext, img: np.ndarray, mw: int, mh: int):
    """load_views' input stage on the device: a decoded u8 image (h, w) or
    (h, w, 3) resized to mw x mh with the Triangle filter, then
    (gray (mh, mw), rgb (mh, mw, 3)) as to_luma8 / to_rgb8 make them."""
    a = np.ascontiguousarray(img, np.uint8)
    ch = 1 if a.ndim == 2 else a.shape[2]
    oh, ow = a.shape[:2]
    gray = np.empty((mh, mw), np.uint8)
    rgb = np.empty((mh, mw, 3), np.uint8)
    check(ctx._lib.mvs_resize_image(ctx._h, ptr(a), int(ch), int(ow), int(oh), int(mw), int(mh), ptr(gray), ptr(rgb)),
          "mvs_resize_image")
    return gray, rgb


def sweep_volume(ctx: MvsContext, job, stride: int = 1) -> np.ndarray:
    """32-plane census sweep volume of a job on the grid of every stride-th
    pixel: (ceil(h/stride), ceil(w/stride), 32) plane costs."""
    arr, keep = _jobs_c([job])
    v = ctx.views[job[0]]
    out = np.empty(coarse_shape(v.h, v.w, stride) + (32,), np.float32)
    check(ctx._lib.mvs_sweep_volume(ctx._h, arr, int(stride), ptr(out)), "mvs_sweep_volume")
    return out


def estimate_depth_learned(ctx: MvsContext, net: ProposalNet, jobs, opts: MvsOptions, vol_k: int = 4,
                           conf_thr: float = 0.5, download: bool = True, extra: bool = True):
    """Learned-init stereo over `jobs` ((ref, srcs, zmin, zmax) tuples).
    extra=True (default): an accepted proposal is a 49th hypothesis of its
    pixel's plain sweep, kept only when its cost is lower; extra=False: it
    replaces the plain sweep at its pixel.
    Returns ({ref: DepthMap} or None, number of pixels with an accepted proposal)."""
    arr, keep = _jobs_c(jobs)
    out = {}
    if download:
        for i, jb in enumerate(jobs):
            v = ctx.views[jb[0]]
            d = np.empty((v.h, v.w), np.float32)
            c = np.empty((v.h, v.w), np.float32)
            n = np.empty((v.h, v.w, 3), np.float32)
            keep += [d, c, n]
            arr[i].depth = d.ctypes.data_as(C.POINTER(C.c_float))
            arr[i].cost = c.ctypes.data_as(C.POINTER(C.c_float))
            arr[i].normals = n.ctypes.data_as(C.POINTER(C.c_float))
            out[int(jb[0])] = DepthMap(d, c, n)
    acc = C.c_int64()
    oc = opts.to_c()
    check(ctx._lib.mvs_estimate_depth_learned_ex(ctx._h, net._h, arr, len(jobs), C.byref(oc), int(vol_k),
                                                 C.c_float(conf_thr), 1 if extra else 0, C.byref(acc)),
          "mvs_estimate_depth_learned_ex")
    return (out if download else None), int(acc.value)


def _tracks_c(xyz, tracks):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    off = np.zeros(len(tracks) + 1, np.int32)
    for i, t in enumerate(tracks):
        off[i + 1] = off[i] + len(t)
    flat = np.asarray([v for t in tracks for v in t] or [0], np.int32)
    return xyz, off, flat


def plan(views: Sequence[ViewData], xyz: np.ndarray, tracks: Sequence[Sequence[int]], opts: MvsOptions):
    """run_mvs's host scheduling (lib.rs:224-282, 326-347, 398-418):
    returns (sources, (zmin, zmax) per view, fusion order, covisibility)."""
    lib = _lib.load()
    n = len(views)
    arr, keep = views_c(views)
    xyz, off, flat = _tracks_c(xyz, tracks)
    so = np.zeros(n + 1, np.int32)
    sr = np.zeros(max(1, n * max(1, int(opts.num_sources))), np.int32)
    zr = np.zeros(2 * n, np.float64)
    fo = np.zeros(n + 1, np.int32)
    fu = np.zeros(n * n, np.int32)
    cov = np.zeros((n, n), np.uint32)
    oc = opts.to_c()
    r = lib.mvs_plan(arr, n, ptr(xyz), ptr(off), ptr(flat), len(xyz), C.byref(oc), ptr(so), ptr(sr), ptr(zr),
                     ptr(fo), ptr(fu), ptr(cov))
    if r != 0:
        raise _lib.MvsError(r, "mvs_plan", "sparse reconstruction too small for MVS" if r == -4 else "bad input")
    sources = [list(sr[so[i]:so[i + 1]]) for i in range(n)]
    fusion = [list(fu[fo[i]:fo[i + 1]]) for i in range(n)]
    return sources, [(float(zr[2 * i]), float(zr[2 * i + 1])) for i in range(n)], fusion, cov


def run_mvs(views: Sequence[ViewData], xyz: np.ndarray, tracks: Sequence[Sequence[int]], out_dir: Optional[str],
            opts: MvsOptions, ctx: Optional[MvsContext] = None,
            log: Optional[Callable[[str], None]] = None) -> dict:
    """run_mvs (lib.rs:294-517) for views at working resolution: plan,
    batched stereo, fusion, <out_dir>/dense/dense.ply, dense_web.json and
    mvs_stats.json, and with opts.save_depth_vis the depth_vis PNGs; `log`
    receives the reference's progress lines.  Returns MvsStats' fields plus
    timings."""
    ctx = ctx or default_context()
    arr, keep = views_c(views)
    xyz, off, flat = _tracks_c(xyz, tracks)
    st = _lib.MvsRunStatsC()
    valid = np.zeros(len(views), np.int64)
    st.per_view_valid = valid.ctypes.data_as(C.POINTER(C.c_int64))
    oc = opts.to_c()
    prev = ctx.set_log(log) if log else None
    try:
        check(ctx._lib.mvs_run_mvs(ctx._h, arr, len(views), ptr(xyz), ptr(off), ptr(flat), len(xyz), C.byref(oc),
                                   None if out_dir is None else str(out_dir).encode(), C.byref(st)), "mvs_run_mvs")
    finally:
        if log:
            ctx.set_log(prev)
    ctx.views = list(views)
    return {"views": st.views, "views_with_depth": st.views_with_depth, "dense_points": st.dense_points,
            "per_view_valid": [int(v) for v in valid],
            "n_candidates": st.n_candidates, "cell": st.cell, "stereo_ms": st.stereo_ms, "fuse_ms": st.fuse_ms}
