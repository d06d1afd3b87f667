"""TEST INFRASTRUCTURE ONLY -- ctypes wrapper of the CPU oracle (mvs_oracle.c).

The oracle is a C restatement of the reference `mvs` crate's dense path
(mvs/src/lib.rs:119-142, mvs/src/stereo.rs, mvs/src/fuse.rs,
sfm-lite/src/geom.rs:27-96).  Only tests, the smoke check and the
benchmark's CPU-baseline leg import this module; the product path
(rustcolmap_amd) never does.

Parity pins of each oracle function are in tests/test_oracle_*.py; the
functions without an independent pin are listed in the "parity unpinned"
section of the oracle's documentation: the multi-pass trajectory of
estimate_depth as a whole (each single pass step is pinned), nalgebra's f64
operation order, and the HashMap output order of fusion.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmvs_oracle.so")


class OrcView(C.Structure):
    _fields_ = [
        ("w", C.c_int32), ("h", C.c_int32), ("f", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k", C.c_double), ("R", C.c_double * 9), ("t", C.c_double * 3),
        ("gray", C.c_void_p), ("rgb", C.c_void_p), ("census", C.c_void_p),
    ]


class OrcOptions(C.Structure):
    _fields_ = [
        ("max_dim", C.c_uint32), ("num_sources", C.c_uint32), ("sweeps", C.c_uint32),
        ("num_hyp", C.c_uint32), ("min_consistent", C.c_uint32), ("save_depth_vis", C.c_uint32),
        ("depth_diff", C.c_double), ("voxel_rel", C.c_double), ("cost_thr", C.c_float),
        ("reserved", C.c_uint32),
    ]


class OrcEvalInfo(C.Structure):
    _fields_ = [("ham_sum", C.c_int32), ("n_invalid", C.c_int32), ("full_sum", C.c_float),
                ("peak", C.c_float), ("aborted", C.c_int32), ("result", C.c_float)]


class OrcPoint(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("pad", C.c_uint8 * 5)]


POINT_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("rgb", "u1", (3,)), ("pad", "u1", (5,))])

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        from . import build
        build.build_oracle()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sigs = {
        "orc_census": (None, [P, C.c_int, C.c_int, P]),
        "orc_pixel_rng": (C.c_uint64, [C.c_int32, C.c_int32, C.c_uint64]),
        "orc_next_u64": (C.c_uint64, [P]),
        "orc_uniform": (C.c_float, [P]),
        "orc_gauss": (C.c_float, [P]),
        "orc_normalize": (None, [C.c_double] * 6 + [P]),
        "orc_project": (C.c_int, [P, P, C.c_double, C.c_double, C.c_double, C.c_double, P, P]),
        "orc_quat_to_rot": (None, [P, P]),
        "orc_match_pair": (C.c_int64, [P, C.c_int, P, C.c_int, C.c_double, P, P]),
        "orc_ws_new": (P, [P, C.c_int, C.c_int, P, C.c_int, C.c_double, C.c_double]),
        "orc_ws_free": (None, [P]),
        "orc_ws_export": (None, [P, P, P, P]),
        "orc_plane_cost": (C.c_float, [P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.POINTER(OrcEvalInfo)]),
        "orc_schedule": (None, [C.c_uint32, C.c_uint32, P, P]),
        "orc_estimate_depth": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int, C.c_double, C.c_double,
                                         C.POINTER(OrcOptions), C.c_int, C.c_int, P, P, P, P, P]),
        "orc_estimate_depth_band": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int, C.c_double, C.c_double,
                                              C.POINTER(OrcOptions), C.c_int, C.c_int, C.c_int, C.c_int,
                                              P, P, P, P, P]),
        "orc_run_passes": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, P, P]),
        "orc_sample_depth": (C.c_double, [P, C.c_int, C.c_int, C.c_double, C.c_double]),
        "orc_fusion_cell": (C.c_double, [P, C.c_int64, C.c_double]),
        "orc_fuse": (C.c_int64, [P, C.c_int, P, P, P, C.c_int, C.c_double, C.c_double, C.c_int,
                                 C.POINTER(C.POINTER(OrcPoint)), C.POINTER(C.POINTER(C.c_double)),
                                 C.POINTER(C.c_int64)]),
        "orc_free": (None, [P]),
        "orc_write_ply": (C.c_int, [C.c_char_p, P, C.c_int64]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def options(opts) -> OrcOptions:
    o = OrcOptions()
    o.max_dim = int(opts.max_dim)
    o.num_sources = int(opts.num_sources)
    o.sweeps = int(opts.sweeps)
    o.num_hyp = int(opts.num_hyp)
    o.min_consistent = int(opts.min_consistent)
    o.save_depth_vis = int(bool(opts.save_depth_vis))
    o.depth_diff = float(opts.depth_diff)
    o.voxel_rel = float(opts.voxel_rel)
    o.cost_thr = float(opts.cost_thr)
    return o


def census(gray: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(gray, np.uint8)
    h, w = g.shape
    out = np.empty((h, w), np.uint64)
    load().orc_census(_p(g), w, h, _p(out))
    return out


class OracleViews:
    """Holds the oracle's view table (census computed by the oracle)."""

    def __init__(self, views, with_census: bool = True):
        """with_census=False skips the census (fusion reads none)."""
        lib = load()
        self.keep = []
        self.rgb_keep = []
        self.arr = (OrcView * len(views))()
        self.views = list(views)
        for i, v in enumerate(views):
            g = np.ascontiguousarray(v.gray, np.uint8)
            cen = census(g) if with_census else np.zeros(1, np.uint64)
            self.keep += [g, cen]   # keep[2i] gray, keep[2i+1] census
            c = self.arr[i]
            c.w, c.h = int(v.w), int(v.h)
            cam = getattr(v, "cam", None)
            if cam is not None:
                c.f, c.cx, c.cy, c.k = float(cam.f), float(cam.cx), float(cam.cy), float(cam.k)
                R = np.asarray(v.pose.r, np.float64).reshape(9)
                t = np.asarray(v.pose.t, np.float64).reshape(3)
            else:
                c.f, c.cx, c.cy, c.k = float(v.f), float(v.cx), float(v.cy), float(v.k)
                R = np.asarray(v.R, np.float64).reshape(9)
                t = np.asarray(v.t, np.float64).reshape(3)
            for j in range(9):
                c.R[j] = float(R[j])
            for j in range(3):
                c.t[j] = float(t[j])
            c.gray = g.ctypes.data
            c.census = cen.ctypes.data if with_census else None
            if v.rgb is not None:
                self.rgb_keep.append(np.ascontiguousarray(v.rgb, np.uint8))
                c.rgb = self.rgb_keep[-1].ctypes.data
            else:
                c.rgb = None
        self.census = [self.keep[2 * i + 1] for i in range(len(views))]
        del lib

    def workspace(self, ref: int, srcs: Sequence[int], zmin: float, zmax: float):
        return Workspace(self, ref, srcs, zmin, zmax)

    def estimate_depth(self, ref, srcs, zmin, zmax, opts, max_passes=-1, nthreads=None, want_planes=False,
                       rows=None):
        """Oracle estimate_depth; `rows=(lo, hi)` computes only that band
        (exact for rows at distance >= passes from band edges)."""
        lib = load()
        v = self.views[ref]
        s = np.ascontiguousarray(srcs, np.int32)
        n = v.w * v.h
        depth = np.empty((v.h, v.w), np.float32)
        cost = np.empty((v.h, v.w), np.float32)
        normals = np.empty((v.h, v.w, 3), np.float32)
        planes = np.empty((v.h, v.w, 5), np.float32) if want_planes else None
        n_passes = 2 * max(int(opts.sweeps), 1)
        n_eval = np.zeros(1 + n_passes, np.int64)
        o = options(opts)
        nt = nthreads or os.cpu_count() or 1
        lo, hi = rows if rows is not None else (0, 0)
        r = lib.orc_estimate_depth_band(C.cast(self.arr, C.c_void_p), len(self.views), int(ref), _p(s), len(s),
                                        float(zmin), float(zmax), C.byref(o), int(max_passes), int(nt), int(lo),
                                        int(hi), _p(depth), _p(cost), _p(normals), _p(planes), _p(n_eval))
        if r != 0:
            raise RuntimeError("orc_estimate_depth failed")
        del n
        return depth, cost, normals, planes, n_eval

    def run_passes(self, ref, srcs, zmin, zmax, planes, pass_lo, pass_hi, n_passes, nthreads=None, variant=0):
        """Passes pass_lo..pass_hi-1 (of n_passes) of stereo.rs:271-371 from
        the given (h, w, 5) planes instead of the init sweep.  Returns the
        planes after the last pass and the evaluation count of each pass.
        variant > 0 runs one of the oracle's planted mistakes (tests only)."""
        lib = load()
        v = self.views[ref]
        s = np.ascontiguousarray(srcs, np.int32)
        pl = np.ascontiguousarray(planes, np.float32).reshape(v.h, v.w, 5).copy()
        ne = np.zeros(max(pass_hi - pass_lo, 1), np.int64)
        r = lib.orc_run_passes(C.cast(self.arr, C.c_void_p), len(self.views), int(ref), _p(s), len(s), float(zmin),
                               float(zmax), int(pass_lo), int(pass_hi), int(n_passes),
                               int(nthreads or os.cpu_count() or 1), int(variant), _p(pl), _p(ne))
        if r != 0:
            raise RuntimeError("orc_run_passes failed")
        return pl, ne[: pass_hi - pass_lo]

    def fuse(self, depth_maps: Sequence[Optional[np.ndarray]], sources, min_consistent, depth_diff, cell,
             nthreads=None, want_candidates=False):
        lib = load()
        nv = len(self.views)
        arr = (C.c_void_p * nv)()
        keep = []
        for i in range(nv):
            d = depth_maps[i]
            if d is None:
                arr[i] = None
            else:
                d = np.ascontiguousarray(d, np.float32)
                keep.append(d)
                arr[i] = d.ctypes.data
        off = np.zeros(nv + 1, np.int32)
        flat = []
        for i in range(nv):
            flat += list(sources[i])
            off[i + 1] = off[i] + len(sources[i])
        flat = np.asarray(flat if flat else [0], np.int32)
        pp = C.POINTER(OrcPoint)()
        cp = C.POINTER(C.c_double)()
        nc = C.c_int64()
        n = lib.orc_fuse(C.cast(self.arr, C.c_void_p), nv, C.cast(arr, C.c_void_p), _p(off), _p(flat),
                         int(min_consistent), float(depth_diff), float(cell), int(nthreads or os.cpu_count() or 1),
                         C.byref(pp), C.byref(cp) if want_candidates else None, C.byref(nc))
        pts = np.zeros(n, POINT_DTYPE)
        if n:
            C.memmove(pts.ctypes.data, pp, n * C.sizeof(OrcPoint))
        lib.orc_free(pp)
        cand = None
        if want_candidates:
            cand = np.ctypeslib.as_array(cp, shape=(max(int(nc.value), 1) * 8,)).copy()[: int(nc.value) * 8]
            cand = cand.reshape(-1, 8)
            lib.orc_free(cp)
        return pts, int(nc.value), cand


class Workspace:
    """RefWorkspace (stereo.rs:103-152) of the oracle, for plane_cost probes."""

    def __init__(self, ov: OracleViews, ref, srcs, zmin, zmax):
        self.lib = load()
        self.ov = ov
        s = np.ascontiguousarray(srcs, np.int32)
        self.k = len(s)
        self.ws = self.lib.orc_ws_new(C.cast(ov.arr, C.c_void_p), len(ov.views), int(ref), _p(s), len(s),
                                      float(zmin), float(zmax))
        if not self.ws:
            raise ValueError("orc_ws_new failed")
        v = ov.views[ref]
        self.w, self.h = v.w, v.h

    def __del__(self):
        try:
            self.lib.orc_ws_free(self.ws)
        except Exception:
            pass

    def export(self):
        rx = np.empty((self.h, self.w), np.float32)
        ry = np.empty((self.h, self.w), np.float32)
        sp = np.empty((self.k, 16), np.float32)
        self.lib.orc_ws_export(self.ws, _p(rx), _p(ry), _p(sp))
        return rx, ry, sp

    def plane_cost(self, px, py, nx, ny, nz, d, hint, info=False):
        if info:
            inf = OrcEvalInfo()
            c = self.lib.orc_plane_cost(self.ws, int(px), int(py), float(nx), float(ny), float(nz), float(d),
                                        float(hint), C.byref(inf))
            return c, inf
        return self.lib.orc_plane_cost(self.ws, int(px), int(py), float(nx), float(ny), float(nz), float(d),
                                       float(hint), None)


def pixel_rng(x, y, salt) -> int:
    return int(load().orc_pixel_rng(int(x), int(y), int(salt)))


def uniform_stream(seed: int, n: int) -> List[float]:
    s = C.c_uint64(seed)
    lib = load()
    return [lib.orc_uniform(C.byref(s)) for _ in range(n)]


def gauss_stream(seed: int, n: int) -> List[float]:
    s = C.c_uint64(seed)
    lib = load()
    return [lib.orc_gauss(C.byref(s)) for _ in range(n)]


def schedule(p: int, n: int):
    a = C.c_float()
    b = C.c_float()
    load().orc_schedule(p, n, C.byref(a), C.byref(b))
    return a.value, b.value


def sample_depth(depth: np.ndarray, u: float, v: float) -> float:
    d = np.ascontiguousarray(depth, np.float32)
    h, w = d.shape
    return float(load().orc_sample_depth(_p(d), w, h, float(u), float(v)))


def fusion_cell(xyz: np.ndarray, voxel_rel: float) -> float:
    a = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    return float(load().orc_fusion_cell(_p(a), len(a), float(voxel_rel)))


def write_ply(path: str, pts: np.ndarray) -> int:
    p = np.ascontiguousarray(pts, POINT_DTYPE)
    return load().orc_write_ply(str(path).encode(), _p(p), len(p))


def normalize(f, cx, cy, k, u, v):
    out = (C.c_double * 2)()
    load().orc_normalize(f, cx, cy, k, u, v, out)
    return out[0], out[1]


def quat_to_rot(q) -> np.ndarray:
    qa = np.ascontiguousarray(q, np.float64)
    R = np.empty(9, np.float64)
    load().orc_quat_to_rot(_p(qa), _p(R))
    return R.reshape(3, 3)


def match_pair(da: np.ndarray, db: np.ndarray, ratio: float = 0.8) -> np.ndarray:
    """(n, 2) int32 index pairs (a, b): mutual nearest neighbours of two
    128-D f32 descriptor sets that pass the symmetric ratio test
    (sfm-lite/src/matching.rs:22-98)."""
    a = np.ascontiguousarray(da, np.float32).reshape(-1, 128)
    b = np.ascontiguousarray(db, np.float32).reshape(-1, 128)
    m = max(min(len(a), len(b)), 1)
    oa = np.empty(m, np.int32)
    ob = np.empty(m, np.int32)
    n = load().orc_match_pair(_p(a), len(a), _p(b), len(b), float(ratio), _p(oa), _p(ob))
    return np.stack([oa[:n], ob[:n]], axis=1)
