"""The C-ABI library: loads, exports every symbol include/mvs_hip.h declares,
and its host-only entry points (defaults, robust cell, PLY writer) agree
with the oracle.  No GPU computation here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from rustcolmap_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mvs_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mvs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers exactly the header
    assert sorted(_lib._SIGS) == names


def test_struct_layouts():
    assert C.sizeof(_lib.MvsOptionsC) == 48
    assert C.sizeof(_lib.MvsPointC) == 32
    assert C.sizeof(_lib.MvsViewC) == 8 + 4 * 8 + 12 * 8 + 24
    assert C.sizeof(_lib.MvsJobC) == 56
    assert C.sizeof(_lib.MvsStatsC) == 11 * 8 + 2 * 4 + 2 * 8 + 4 * 4


def test_options_default_matches_reference():
    lib = _lib.load()
    o = _lib.MvsOptionsC()
    lib.mvs_options_default(C.byref(o))
    assert (o.max_dim, o.num_sources, o.sweeps, o.num_hyp, o.min_consistent) == (900, 3, 4, 48, 2)
    assert o.depth_diff == 0.03 and o.voxel_rel == 0.0015 and o.cost_thr == np.float32(0.38)
    from rustcolmap_amd.mvs import MvsOptions
    d = MvsOptions()
    assert (d.max_dim, d.num_sources, d.sweeps, d.num_hyp, d.min_consistent) == (900, 3, 4, 48, 2)


def test_fusion_cell_matches_oracle():
    from rustcolmap_amd.mvs import fusion_cell
    rng = np.random.default_rng(0)
    for n in (1, 7, 100, 5000):
        xyz = rng.normal(size=(n, 3)) * [3, 1, 0.5]
        assert fusion_cell(xyz, 0.0015) == oracle.fusion_cell(xyz, 0.0015)
    with pytest.raises(ValueError):
        fusion_cell(np.zeros((0, 3)), 0.0015)


def test_write_ply_matches_oracle(tmp_path):
    from rustcolmap_amd.mvs import POINT_DTYPE, write_ply_binary
    rng = np.random.default_rng(1)
    pts = np.zeros(1000, POINT_DTYPE)
    pts["x"], pts["y"], pts["z"] = rng.normal(size=(3, 1000))
    pts["rgb"] = rng.integers(0, 256, size=(1000, 3))
    a, b = tmp_path / "a.ply", tmp_path / "b.ply"
    write_ply_binary(str(a), pts)
    oracle.write_ply(str(b), pts)
    assert a.read_bytes() == b.read_bytes()
    write_ply_binary(str(a), pts[:0])
    oracle.write_ply(str(b), pts[:0])
    assert a.read_bytes() == b.read_bytes()


def test_errors_are_reported_not_raised_across_the_boundary():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.mvs_ctx_create(0, None) == -1
    assert b"NULL" in lib.mvs_last_error()
    assert lib.mvs_write_ply(b"/nonexistent-dir/x.ply", None, 0) != 0
    assert lib.mvs_fusion_cell(None, 0, 0.1) < 0
    del h
