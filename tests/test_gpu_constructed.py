"""GPU parity on constructed pass states and at K = 12.

* The constructed scenes of test_oracle_pass_pins (an exact integer-
  disparity source; planes given instead of the init sweep: one zero-cost
  seed pixel, a uniform plane at the inv_new rejection boundary, and a
  checkerboard of a near-band and a far plane) run through the GPU's pass
  kernels (mvs_debug_run_passes) in several launch families; planes and
  evaluation counts must equal the oracle's (whose answers on these scenes
  the geometry fixes) bit for bit.
* 13 views of 160x120 with a different radial k per view (some 0): every
  view uses the 12 others as sources (K = 12, the largest K of the uneven
  batch), planes after every pass and evaluation counts against the oracle.
"""
import dataclasses

import numpy as np
import pytest

import test_oracle_pass_pins as pins
from conftest import bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _view_data(ns):
    from rustcolmap_amd.mvs import Cam, Pose, ViewData
    return ViewData(w=ns.w, h=ns.h, cam=Cam(ns.f, ns.cx, ns.cy, ns.k), pose=Pose(np.asarray(ns.R, np.float64),
                                                                                  np.asarray(ns.t, np.float64)),
                    gray=ns.gray)


def _states():
    """(name, initial planes, pass index) of the constructed cases."""
    out = []
    p0 = pins._planes(pins._fronto(pins.M + 5))
    p0[12, 24] = pins._fronto(pins.M, cost=0.0)
    out.append(("ring", p0, 0))
    for p in pins.PASSES:
        z0 = np.float32(1.0 / (0.05 + 0.3 * pins._reach(p) / 1.5))
        out.append(("reject%d" % p, pins._planes(pins._fronto(pins.FB / z0)), p))
        out.append(("checker%d" % p, pins._checkerboard(p)[0], p))
    return out


@pytest.mark.parametrize("shape", ["8x8pw", "16x16", "16x8r", "8x8v", "32x8w", "16x16px", "16x16sw", "24x8sw"])
def test_constructed_pass_states(ctx, shape, monkeypatch):
    monkeypatch.setenv("MVS_TILE", shape)
    ov = pins._scene()
    ctx.upload_views([_view_data(v) for v in ov.views])
    for name, planes, p in _states():
        _say("constructed %s pass %d shape %s" % (name, p, shape))
        n = 4 if name == "ring" else 1
        got, ne = ctx.debug_run_passes(0, [1], pins.ZMIN, pins.ZMAX, planes, p, p + n, pins.N_PASSES)
        exp, ne_o = ov.run_passes(0, [1], pins.ZMIN, pins.ZMAX, planes, p, p + n, pins.N_PASSES)
        assert np.array_equal(bits(got), bits(exp)), (name, np.argwhere(bits(got) != bits(exp))[:4])
        assert np.array_equal(ne, ne_o), (name, ne, ne_o)
    # the ring after 4 passes is the Manhattan ball of radius 4
    star = pins._fronto(pins.M, cost=0.0)
    yy, xx = np.mgrid[0:pins.H, 0:pins.W]
    ring_planes, _ = ctx.debug_run_passes(0, [1], pins.ZMIN, pins.ZMAX, _states()[0][1], 0, 4, pins.N_PASSES)
    assert np.array_equal(pins._is(ring_planes, star), np.abs(xx - 24) + np.abs(yy - 12) <= 4)


KS = [0.05, 0.0, -0.04, 0.08, 0.0, -0.08, 0.02, 0.11, 0.0, -0.02, 0.06, -0.06, 0.03]


def _k12_views():
    from rustcolmap_amd.mvs import ViewData
    from rustcolmap_amd.scene import SceneSpec, render_scene
    sc = render_scene(SceneSpec(seed=31, n_views=13, sizes=((160, 120),), focal=140.0, arc_deg=30.0, radius=4.0,
                                arc_center=-0.5, n_objects=3, tex_freq=6.0), device="cpu")
    views = [ViewData.from_scene_view(v) for v in sc.views]
    return [dataclasses.replace(v, cam=dataclasses.replace(v.cam, k=KS[i])) for i, v in enumerate(views)]


@pytest.mark.parametrize("shape", ["8x8pw", "16x16", "16x8v", "32x8w", "16x8sw"])
def test_k12_every_pass(ctx, shape, monkeypatch):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    monkeypatch.setenv("MVS_TILE", shape)
    views = _k12_views()
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=12, sweeps=2)
    for ref in (0, 6, 12):
        srcs = [j for j in range(13) if j != ref]
        for mp in (0, 1, 2, 4):
            _say("K=12 ref %d passes %d shape %s" % (ref, mp, shape))
            dm, planes, ne = ctx.estimate_depth(ref, srcs, 1.0, 10.0, opts, max_passes=mp, want_planes=True)
            d, c, n, pl, ne_o = ov.estimate_depth(ref, srcs, 1.0, 10.0, opts, max_passes=mp, want_planes=True)
            assert np.array_equal(bits(planes), bits(pl)), (ref, mp, np.argwhere(bits(planes) != bits(pl))[:4])
            assert np.array_equal(bits(dm.depth), bits(d)), (ref, mp)
            assert np.array_equal(bits(dm.cost), bits(c)), (ref, mp)
            assert np.array_equal(ne[: 1 + mp], ne_o[: 1 + mp]), (ref, mp, ne, ne_o)
