"""GPU parity of the HIP path against the CPU oracle (bit-exact).

Integer, index and float work of the stereo path is compared bit for bit:
census codes, RNG streams, per-evaluation Hamming sums, invalid-tap counts,
returned costs, plane buffers after every pass, gated depth / cost /
normals and evaluation counts.  Fusion output is compared point by point
(same first-insertion order on both sides).
"""
import numpy as np
import pytest

from conftest import bits, scene, view_data

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_census_bitexact(ctx):
    import oracle
    rng = np.random.default_rng(5)
    for (h, w) in [(1, 1), (2, 3), (7, 7), (8, 33), (31, 64), (240, 320), (97, 131)]:
        _say("census %dx%d" % (w, h))
        g = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
        if h * w > 4:
            g[: h // 2, : w // 2] = 77      # flat region
        got = ctx.census(g)
        exp = oracle.census(g)
        assert np.array_equal(got, exp), (h, w)


def test_view_census_and_upload(ctx):
    import oracle
    sc = scene("tiny")
    views = view_data(sc)
    ctx.upload_views(views)
    for i, v in enumerate(views):
        assert np.array_equal(ctx.view_census(i), oracle.census(v.gray))


def test_cvt_matches_rust_as_i32(ctx):
    """v_cvt_i32_f32 as used by the stereo kernels: truncate toward zero,
    saturate, NaN -> 0 -- Rust's `f32 as i32`."""
    import ctypes as C
    from rustcolmap_amd import _lib
    xs = np.array([0.0, -0.0, 0.5, -0.5, -0.999, -1.0, -1.5, 1.999, 2147483520.0, 2147483648.0, 3e9, -2147483648.0,
                   -3e9, np.inf, -np.inf, np.nan, -np.nan, 1e-45, -1e-45, 123456.7, -123456.7, 1919.9999,
                   np.float32(np.uint32(0x7fc00001).view(np.float32))], np.float32)
    out = np.empty(len(xs), np.int32)
    _lib.check(_lib.load().mvs_debug_cvt(ctx._h, xs.ctypes.data_as(C.c_void_p), len(xs),
                                         out.ctypes.data_as(C.c_void_p)), "mvs_debug_cvt")

    def rust(x):
        if np.isnan(x):
            return 0
        if x >= 2147483648.0:
            return 2147483647
        if x <= -2147483648.0:
            return -2147483648
        return int(np.trunc(x))
    for x, o in zip(xs, out):
        assert int(o) == rust(float(x)), (x, o)


def test_fast_divide_is_ieee_division(ctx):
    """The stereo kernels' scaling-free division equals IEEE f32 division
    bit for bit over its domain (b and 1/b normal, a / b normal,
    |a| >= 2^-104): log-uniform operands over the ranges the taps produce and
    beyond, exact multiples, and quotients one product away from
    representable numbers."""
    import ctypes as C
    from rustcolmap_amd import _lib
    rng = np.random.default_rng(11)
    n = 1 << 22
    parts = []
    for lo_a, hi_a, lo_b, hi_b in ((-104, 60, -14, 60), (-20, 45, -10, 10), (-3, 3, -5, 1)):
        a = np.exp2(rng.uniform(lo_a, hi_a, n)) * rng.choice([-1.0, 1.0], n)
        b = np.exp2(rng.uniform(lo_b, hi_b, n))
        parts.append((a.astype(np.float32), b.astype(np.float32)))
    q = (rng.uniform(1, 2, n) * np.exp2(rng.integers(-30, 30, n))).astype(np.float32)
    b = (rng.uniform(0.05, 1000.0, n)).astype(np.float32)
    parts.append(((q * b).astype(np.float32), b))
    ones = np.float32(np.uint32(0x3fffffff).view(np.float32))     # all-ones significand
    parts.append((rng.uniform(-1e6, 1e6, 1024).astype(np.float32), np.full(1024, ones, np.float32)))
    a = np.concatenate([p[0] for p in parts])
    b = np.concatenate([p[1] for p in parts])
    with np.errstate(all="ignore"):
        ref = a / b
    keep = (np.abs(a) >= np.float32(2.0 ** -104)) & (np.abs(ref) >= np.float32(2.0 ** -126)) & \
           (np.abs(ref) < np.float32(2.0 ** 96)) & (b >= np.float32(2.0 ** -126)) & (b < np.float32(2.0 ** 126))
    a, b, ref = a[keep], b[keep], ref[keep]
    out = np.empty_like(a)
    _lib.check(_lib.load().mvs_debug_div(ctx._h, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), len(a),
                                         out.ctypes.data_as(C.c_void_p)), "mvs_debug_div")
    bad = np.flatnonzero(out.view(np.uint32) != ref.view(np.uint32))
    _say("fast divide: %d quotients, %d differ" % (len(a), len(bad)))
    assert len(bad) == 0, [(float(a[i]), float(b[i]), float(out[i]), float(ref[i])) for i in bad[:5]]


def test_rng_streams(ctx):
    import oracle
    rng = np.random.default_rng(1)
    xy = np.concatenate([rng.integers(-5, 3000, size=(200, 2)), [[0, 0], [-1, -1], [2**31 - 1, -2**31]]])
    salt = np.concatenate([rng.integers(0, 40, size=200), [7, 0, 2**63 + 5]]).astype(np.uint64)
    seed, uni = ctx.debug_rng(xy, salt, 12)
    for t in range(len(xy)):
        s = oracle.pixel_rng(int(xy[t, 0]), int(xy[t, 1]), int(salt[t]))
        assert int(seed[t]) == s
        assert np.array_equal(bits(uni[t]), bits(np.asarray(oracle.uniform_stream(s, 12), np.float32)))


def _queries(rng, w, h, n, zmin=1.0, zmax=10.0):
    q = np.zeros((n, 7), np.float32)
    q[:, 0] = rng.integers(0, w, n)
    q[:, 1] = rng.integers(0, h, n)
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = np.abs(nrm[:, 2]) + 0.3
    nrm[: n // 3] = [0.0, 0.0, 1.0]
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    q[:, 2:5] = nrm
    q[:, 5] = rng.uniform(zmin * 0.4, zmax * 2.5, n)
    hint = rng.choice([np.inf, 0.1, 0.25, 0.4, 1.02], n)
    q[:, 6] = hint
    # a few planes behind the camera / grazing
    q[-5:, 4] = -0.2
    q[-10:-5, 5] = -1.0
    return q


def test_plane_cost_parity(ctx):
    import oracle
    for name, k in (("tiny", 2), ("tiny", 4), ("corner320", 4)):
        sc = scene(name)
        views = view_data(sc)
        ctx.upload_views(views)
        ov = oracle.OracleViews(views)
        rng = np.random.default_rng(k)
        from rustcolmap_amd.scene import select_sources
        srcs_all = select_sources(sc.views, k)
        for ref in (0, len(views) // 2):
            _say("plane_cost %s ref %d K %d" % (name, ref, k))
            srcs = srcs_all[ref]
            v = views[ref]
            q = _queries(rng, v.w, v.h, 400)
            cost, ham, ninv, peak = ctx.debug_plane_cost(ref, srcs, q)
            ws = ov.workspace(ref, srcs, 1.0, 10.0)
            for t in range(len(q)):
                c, info = ws.plane_cost(*q[t, :2].astype(int), *q[t, 2:7], info=True)
                c0 = ws.plane_cost(*q[t, :2].astype(int), *q[t, 2:7], info=False)
                assert np.float32(c).view(np.uint32) == np.float32(c0).view(np.uint32)
                assert np.float32(cost[t]).view(np.uint32) == np.float32(c).view(np.uint32), (t, cost[t], c)
                assert ham[t] == info.ham_sum and ninv[t] == info.n_invalid, t
                assert np.float32(peak[t]).view(np.uint32) == np.float32(info.peak).view(np.uint32), t


def _run_both(ctx, ov, ref, srcs, opts, max_passes, zmin=1.0, zmax=10.0):
    dm, planes, ne = ctx.estimate_depth(ref, srcs, zmin, zmax, opts, max_passes=max_passes, want_planes=True)
    d, c, n, pl, ne_o = ov.estimate_depth(ref, srcs, zmin, zmax, opts, max_passes=max_passes, want_planes=True)
    return (dm, planes, ne), (d, c, n, pl, ne_o)


def test_estimate_depth_every_pass_tiny(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("tiny")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(sc.views, 3)
    opts = MvsOptions(num_sources=3, sweeps=3)
    for ref in range(len(views)):
        for mp in range(0, 7):
            _say("tiny ref %d passes %d" % (ref, mp))
            (dm, planes, ne), (d, c, n, pl, ne_o) = _run_both(ctx, ov, ref, srcs[ref], opts, mp)
            assert np.array_equal(bits(planes), bits(pl)), (ref, mp, np.argwhere(bits(planes) != bits(pl))[:5])
            assert np.array_equal(bits(dm.depth), bits(d))
            assert np.array_equal(bits(dm.cost), bits(c))
            assert np.array_equal(bits(dm.normals), bits(n))
            assert np.array_equal(ne[: 1 + mp], ne_o[: 1 + mp])


def test_estimate_depth_corner320_all_views(ctx):
    """10 views x 320x240, K=4, 6 sweeps, inverse depth [0.1, 1]."""
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(sc.views, 4)
    opts = MvsOptions(num_sources=4, sweeps=6)
    for ref in range(len(views)):
        _say("corner320 ref %d" % ref)
        (dm, planes, ne), (d, c, n, pl, ne_o) = _run_both(ctx, ov, ref, srcs[ref], opts, -1)
        assert np.array_equal(bits(planes), bits(pl)), ref
        assert np.array_equal(bits(dm.depth), bits(d)), ref
        assert np.array_equal(bits(dm.cost), bits(c)), ref
        assert np.array_equal(ne, ne_o), ref


def test_batch_modes_match_single(ctx):
    """All batch schedules (per-job launches, tile grid, persistent queue,
    graph replay and direct launches) give the single-view results."""
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    # uneven K per job
    ks = [2, 3, 4, 5, 2, 4, 3, 5, 4, 2]
    jobs = []
    for ref in range(len(views)):
        jobs.append((ref, select_sources(sc.views, ks[ref])[ref], 1.0, 10.0))
    exp = {}
    for ref, s, zmin, zmax in jobs:
        d, c, n, _, _ = ov.estimate_depth(ref, s, zmin, zmax, opts)
        exp[ref] = (d, c, n)
    for mode in (1, 2, 3, 6, 7, 2, 3):
        _say("batch mode %d" % mode)
        out = ctx.estimate_depth_batch(jobs, opts, mode=mode)
        for ref in exp:
            assert np.array_equal(bits(out[ref].depth), bits(exp[ref][0])), (mode, ref)
            assert np.array_equal(bits(out[ref].cost), bits(exp[ref][1])), (mode, ref)
            assert np.array_equal(bits(out[ref].normals), bits(exp[ref][2])), (mode, ref)
        # resident copy equals the downloaded one
        dm = ctx.get_depth(3)
        assert np.array_equal(bits(dm.depth), bits(exp[3][0]))


def test_graph_replays_with_eager_work_between(ctx):
    import oracle
    import torch
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("tiny")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    srcs = select_sources(sc.views, 2)
    opts = MvsOptions(num_sources=2, sweeps=2)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    exp = {r: ov.estimate_depth(r, srcs[r], 1.0, 10.0, opts)[0] for r in range(len(views))}
    for rep in range(4):
        _say("graph replay %d" % rep)
        x = torch.randn(1 << 20, device="cuda")
        _ = (x * 2).sum().item()
        torch.cuda.synchronize()
        out = ctx.estimate_depth_batch(jobs, opts, mode=2)
        st = ctx.stats()
        assert st.graph_used == 1
        for r in exp:
            assert np.array_equal(bits(out[r].depth), bits(exp[r])), (rep, r)


def test_fusion_parity(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(sc.views, 4)
    dms = []
    for ref in range(len(views)):
        d, c, n, _, _ = ov.estimate_depth(ref, srcs[ref], 1.0, 10.0, opts)
        dms.append(d)
    dms[7] = None     # a view without a depth map
    order = [[j for j in select_sources(sc.views, 9)[i] if dms[j] is not None] for i in range(len(views))]
    from rustcolmap_amd.mvs import DepthMap
    for cell, mc in ((0.01, 2), (0.004, 1), (0.05, 3)):
        _say("fusion cell %g min_consistent %d" % (cell, mc))
        o2 = MvsOptions(min_consistent=mc)
        got = ctx.fuse_depth_maps([None if d is None else DepthMap(d, d) for d in dms], order, cell, o2)
        exp, ncand, _ = ov.fuse(dms, order, mc, o2.depth_diff, cell)
        assert ctx.stats().n_candidates == ncand
        assert len(got) == len(exp)
        for f in ("x", "y", "z"):
            assert np.array_equal(bits(got[f]), bits(exp[f])), f
        assert np.array_equal(got["rgb"], exp["rgb"])


def test_fusion_resident_matches_host_maps(ctx):
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(sc.views, 4)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    order = select_sources(sc.views, 9)
    a = ctx.fuse_resident([True] * len(views), order, 0.01, opts)
    b = ctx.fuse_depth_maps([out[r] for r in range(len(views))], order, 0.01, opts)
    assert len(a) == len(b) and len(a) > 1000
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", ["16x16", "16x16r", "32x8", "32x8p", "16x8r", "16x8p", "8x8", "16x16pr", "8x8p",
                                   "16x16v", "8x8v", "32x8pv", "16x8v", "16x16w", "16x16x", "16x8w", "16x8px",
                                   "8x8w", "8x8px", "8x8pw", "32x8w", "16x16pw", "16x8sw", "16x16sw", "24x8sw"])
def test_every_launch_shape_is_exact(ctx, shape, monkeypatch):
    """Tile shape, evaluator and schedule change only speed, never values."""
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    monkeypatch.setenv("MVS_TILE", shape)
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(sc.views, 4)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in (0, 4, 9)]
    _say("launch shape %s" % shape)
    for mode in (0, 4):
        out = ctx.estimate_depth_batch(jobs, opts, mode=mode)
        for r, s, zmin, zmax in jobs:
            d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
            assert np.array_equal(bits(out[r].depth), bits(d)), (shape, mode, r)
            assert np.array_equal(bits(out[r].cost), bits(c)), (shape, mode, r)


def test_autotune_then_exact(ctx):
    import oracle
    from rustcolmap_amd.mvs import MvsOptions
    from rustcolmap_amd.scene import select_sources
    sc = scene("corner320")
    views = view_data(sc)
    ctx.upload_views(views)
    ov = oracle.OracleViews(views)
    opts = MvsOptions(num_sources=4, sweeps=2)
    srcs = select_sources(sc.views, 4)
    jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
    best = ctx.autotune(jobs, opts, passes=1)
    _say("autotune picked %s" % (best,))
    assert best[0] * best[1] in (64, 128, 256) and best[2] in (0, 1) and best[3] in (0, 1, 2, 3, 4)
    out = ctx.estimate_depth_batch(jobs, opts, mode=0)
    for r, s, zmin, zmax in jobs:
        d, c, n, _, _ = ov.estimate_depth(r, s, zmin, zmax, opts)
        assert np.array_equal(bits(out[r].depth), bits(d)), r


def test_run_mvs_matches_oracle_pipeline(ctx, tmp_path):
    """run_mvs end to end (plan from a sparse model, batched stereo, fusion,
    dense.ply) equals the oracle run on the same plan."""
    import oracle
    from rustcolmap_amd.mvs import MvsOptions, plan, run_mvs
    from rustcolmap_amd.scene import sparse_points
    sc = scene("corner320")
    views = view_data(sc)
    xyz, tracks = sparse_points(sc, 400, seed=1)
    opts = MvsOptions(num_sources=4, sweeps=2, save_depth_vis=True)
    lines = []
    st = run_mvs(views, xyz, tracks, str(tmp_path), opts, ctx=ctx, log=lines.append)
    src, ranges, fus, _ = plan(views, xyz, tracks, opts)
    ov = oracle.OracleViews(views)
    dms = []
    for i in range(len(views)):
        if not src[i]:
            dms.append(None)
            continue
        d, _, _, _, _ = ov.estimate_depth(i, src[i], ranges[i][0], ranges[i][1], opts)
        dms.append(d)
        assert st["per_view_valid"][i] == int((d > 0).sum()), i
    cell = oracle.fusion_cell(xyz, opts.voxel_rel)
    assert st["cell"] == cell
    exp, ncand, _ = ov.fuse(dms, fus, opts.min_consistent, opts.depth_diff, cell)
    assert st["dense_points"] == len(exp) and st["n_candidates"] == ncand
    ply = (tmp_path / "dense" / "dense.ply").read_bytes()
    ref = tmp_path / "ref.ply"
    oracle.write_ply(str(ref), exp)
    assert ply == ref.read_bytes()
    assert (tmp_path / "dense" / "mvs_stats.json").exists()
    # dense_web.json written by the library = the restatement of
    # mvs/src/lib.rs:437-485 on the same (oracle) points
    from rustcolmap_amd import formats
    from rustcolmap_amd.mvs import POINT_DTYPE
    pts = np.zeros(len(exp), POINT_DTYPE)
    for k in ("x", "y", "z"):
        pts[k] = exp[k]
    pts["rgb"] = exp["rgb"]
    web = (tmp_path / "dense" / "dense_web.json").read_text()
    assert web == formats.dense_web_json(pts)
    import json
    assert json.loads(web)["n"] <= len(exp)
    # save_depth_vis (lib.rs:383-392): <stem>_depth.png of every depth map,
    # the pixels of fuse.rs:243-260 for the (oracle-equal) depth map
    for i, d in enumerate(dms):
        png = tmp_path / "dense" / "depth_vis" / ("v%03d_depth.png" % i)
        assert png.exists() == (d is not None), i
        if d is not None:
            assert np.array_equal(formats.read_png_gray8(png.read_bytes()), formats.depth_vis(d)), i
    # the reference's log lines (lib.rs:212-216, 344-347, 362-371, 376-380,
    # 391, 509-515; fuse.rs:89-92, 170-174), views in index order
    import re
    dbg = formats.rust_f64_debug
    with_src = [i for i in range(len(views)) if src[i]]
    want = [re.escape("mvs: loaded %d views at working resolution (max_dim 900)" % len(views)),
            re.escape("mvs: %d views with >=1 source view (K = 4)" % len(with_src))]
    want += [re.escape("mvs: [v%03d] depth %s-%s (%d srcs) valid %d/%d px in " % (
        i, dbg(ranges[i][0]), dbg(ranges[i][1]), len(src[i]), int((dms[i] > 0).sum()), views[i].w * views[i].h))
        + r"\d+\.\ds" for i in with_src]
    want += [r"mvs: stereo done in \d+\.\ds \(%d depth maps\)" % len(with_src),
             re.escape("mvs: depth visualizations saved")]
    ext = [np.sort(xyz[:, a])[[int(len(xyz) * 0.02), min(int(len(xyz) * 0.98), len(xyz) - 1)]] for a in range(3)]
    diag = float(np.sqrt(sum((hi - lo) ** 2 for lo, hi in ext)))
    want += [re.escape("mvs: fusion voxel cell %.4f (robust sparse diag %.2f)" % (cell, diag)),
             re.escape("mvs: %d consistency-passed candidates from %d depth maps" % (ncand, len(with_src))),
             re.escape("mvs: fused %d dense points -> %s/dense/dense.ply in " % (len(exp), tmp_path))
             + r"\d+\.\ds \(total \d+\.\ds\)"]
    assert len(lines) == len(want), lines
    for got, pat in zip(lines, want):
        assert re.fullmatch(pat, got), (got, pat)


_UNEVEN = {}


def _uneven_case():
    """Six views of three odd sizes (partial tiles on both axes; sources of
    other sizes than the reference), with the oracle's depth and cost."""
    if not _UNEVEN:
        import oracle
        from rustcolmap_amd.mvs import MvsOptions, ViewData
        from rustcolmap_amd.scene import SceneSpec, render_scene, select_sources
        sc = render_scene(SceneSpec(seed=21, n_views=6, sizes=((317, 203), (250, 190), (333, 251)),
                                    focal_rel=0.86, arc_deg=50.0, radius=4.0, arc_center=-0.5,
                                    n_objects=2, tex_freq=8.0), device="cpu")
        views = [ViewData.from_scene_view(v) for v in sc.views]
        ov = oracle.OracleViews(views)
        opts = MvsOptions(num_sources=3, sweeps=2)
        srcs = select_sources(sc.views, 3)
        jobs = [(r, srcs[r], 1.0, 10.0) for r in range(len(views))]
        exp = {r: ov.estimate_depth(r, s, zmin, zmax, opts)[:2] for r, s, zmin, zmax in jobs}
        _UNEVEN.update(views=views, opts=opts, jobs=jobs, exp=exp)
    return _UNEVEN


@pytest.mark.parametrize("shape", ["8x8pw", "16x16w", "32x8w", "16x8r", "16x16sw", "24x8sw"])
def test_uneven_sizes_batch_is_exact(ctx, shape, monkeypatch):
    """One batch over views of different, odd sizes in several launch
    families, single-pass and persistent schedules alike."""
    monkeypatch.setenv("MVS_TILE", shape)
    case = _uneven_case()
    views = case["views"]
    ctx.upload_views(views)
    _say("uneven sizes, launch shape %s" % shape)
    for mode in (0, 4):
        out = ctx.estimate_depth_batch(case["jobs"], case["opts"], mode=mode)
        for r, _, _, _ in case["jobs"]:
            d, c = case["exp"][r]
            assert np.array_equal(bits(out[r].depth), bits(d)), (shape, mode, r, views[r].w, views[r].h)
            assert np.array_equal(bits(out[r].cost), bits(c)), (shape, mode, r)
