"""run_mvs host scheduling (mvs_plan, mvs/src/lib.rs:224-282, 326-347,
398-418) on constructed reconstructions with known answers."""
import numpy as np
import pytest

from rustcolmap_amd import _lib
from rustcolmap_amd.mvs import Cam, MvsOptions, Pose, ViewData, plan


def _views(n):
    out = []
    for i in range(n):
        R = np.eye(3)
        t = np.array([-0.1 * i, 0.0, 0.0])       # cameras along x, looking down +z
        out.append(ViewData(w=8, h=6, cam=Cam(10.0, 4.0, 3.0), pose=Pose(R, t), gray=np.zeros((6, 8), np.uint8)))
    return out


def _points(spec, depth=5.0):
    """spec: list of (track, count): `count` points at `depth` seen by `track`."""
    xyz, tracks = [], []
    for tr, cnt in spec:
        for k in range(cnt):
            xyz.append([0.01 * k, 0.0, depth])
            tracks.append(list(tr))
    return np.asarray(xyz), tracks


def test_sources_top_k_with_threshold_and_ties():
    views = _views(5)
    # shared counts with view 0: v1 30, v2 12, v3 12, v4 9 (below 10)
    xyz, tracks = _points([((0, 1), 18), ((0, 1, 2), 12), ((0, 3), 12), ((0, 4), 9), ((2, 3, 4), 20)])
    src, rng, fus, cov = plan(views, xyz, tracks, MvsOptions(num_sources=3))
    assert cov[0, 1] == 30 and cov[0, 2] == 12 and cov[0, 3] == 12 and cov[0, 4] == 9
    assert src[0] == [1, 2, 3]                      # 30, then the 12-12 tie in index order
    assert src[4] == [2, 3]                         # 29, 29 with view 2/3; view 0 shares 9
    src2, _, _, _ = plan(views, xyz, tracks, MvsOptions(num_sources=1))
    assert src2[0] == [1]


def test_fusion_order_puts_sources_first():
    views = _views(4)
    xyz, tracks = _points([((0, 1), 10), ((0, 2), 40), ((0, 3), 25), ((1, 2, 3), 30)])
    src, _, fus, cov = plan(views, xyz, tracks, MvsOptions(num_sources=2))
    assert src[0] == [2, 3]
    # co-visible views with depth maps by count: 2 (40), 3 (25), 1 (10); sources (2, 3) first, in source order
    assert fus[0] == [2, 3, 1]
    srcb, _, fusb, _ = plan(views, xyz, tracks, MvsOptions(num_sources=1))
    assert srcb[0] == [2] and fusb[0] == [2, 3, 1]


def test_depth_range_percentiles_and_fallback():
    views = _views(3)
    depths = np.linspace(2.0, 6.0, 101)
    xyz = np.stack([np.zeros_like(depths), np.zeros_like(depths), depths], 1)
    tracks = [[0, 1]] * 101
    _, rng, _, _ = plan(views, xyz, tracks, MvsOptions())
    # view 0: sorted depths, idx round(100 * 0.02) = 2 and round(100 * 0.98) = 98
    assert rng[0] == (depths[2] * 0.8, depths[98] * 1.25)
    # view 2 sees no point: falls back to the global |p| percentiles (same depths here)
    assert rng[2] == (depths[2] * 0.8, depths[98] * 1.25)
    # degenerate span (hi <= 1.05 lo) -> last-resort global 1/99 % * 0.7 / 1.4
    # (view-0 depth constant 4, while |p| varies with x)
    xyz2 = np.zeros((20, 3))
    xyz2[:, 2] = 4.0
    xyz2[:, 0] = np.arange(20.0)
    _, rng2, _, _ = plan(views, xyz2, [[0, 1]] * 20, MvsOptions())
    g = np.sort(np.linalg.norm(xyz2, axis=1))
    assert rng2[0] == (max(g[int(round(19 * 0.01))] * 0.7, 1e-3), g[int(round(19 * 0.99))] * 1.4)


def test_too_few_points_is_an_error():
    views = _views(2)
    xyz, tracks = _points([((0, 1), 7)])
    with pytest.raises(_lib.MvsError):
        plan(views, xyz, tracks, MvsOptions())
