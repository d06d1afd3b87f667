"""The pass kernels' skipping of repeated neighbour planes, pinned on the
oracle (CPU).

The device counts a neighbour candidate as an evaluation without running it
when its (n, d_plane) bits equal those of the current plane -- if that plane
is fronto-parallel, (+0, +0, 1, z) -- or else of the first neighbour
evaluated at the pixel (stereo.rs:293-321).  The argument: plane_cost depends
on (n, d_plane) and the hint only, and the best cost only falls, so a repeat
aborts wherever the earlier evaluation did or returns the same cost, which
cannot beat a best that is already at most that cost.  For the current plane
this needs its stored cost to be at most its own cost here, which holds for
every state the init sweep and the passes make, and not for planted states.

The oracle runs the same rule as pass variants 5 (repeats counted) and 6
(not counted); these tests check, on rendered scenes and all their passes,
that variant 5 leaves every plane, cost and evaluation count of the reference
pass unchanged while variant 6 shows how much is skipped, and that the
current-plane rule does change planted states whose stored costs are not the
planes' own -- which is why the device applies it only to states it made.
"""
import numpy as np
import pytest

import oracle
import test_oracle_pass_pins as pins
from conftest import bits, scene, view_data

from rustcolmap_amd.mvs import MvsOptions
from rustcolmap_amd.scene import select_sources

N_PASSES = 6


def _init_planes(ov, ref, srcs, opts):
    _, _, _, planes, _ = ov.estimate_depth(ref, srcs, 1.0, 10.0, opts, max_passes=0, want_planes=True)
    return planes


@pytest.mark.parametrize("ref", [0, 4])
def test_skipping_repeats_changes_no_plane_and_no_count(ref):
    views = view_data(scene("corner320"))
    ov = oracle.OracleViews(views)
    srcs = select_sources(scene("corner320").views, 3)[ref]
    opts = MvsOptions(num_sources=3, sweeps=N_PASSES // 2)
    planes = _init_planes(ov, ref, srcs, opts)
    ref_pl, ref_ne = ov.run_passes(ref, srcs, 1.0, 10.0, planes, 0, N_PASSES, N_PASSES, variant=0)
    skip_pl, skip_ne = ov.run_passes(ref, srcs, 1.0, 10.0, planes, 0, N_PASSES, N_PASSES, variant=5)
    run_pl, run_ne = ov.run_passes(ref, srcs, 1.0, 10.0, planes, 0, N_PASSES, N_PASSES, variant=6)
    assert np.array_equal(bits(skip_pl), bits(ref_pl))
    assert np.array_equal(bits(run_pl), bits(ref_pl))
    assert np.array_equal(skip_ne, ref_ne)
    # the rule is exercised: a large share of the reference's evaluations are repeats
    skipped = 1.0 - run_ne.sum() / ref_ne.sum()
    assert skipped > 0.2, skipped
    # and the planes do change over the passes (the run is not trivially static)
    assert not np.array_equal(bits(ref_pl), bits(planes))


def test_skipping_repeats_from_a_mid_run_state():
    """States after a few reference passes (planes accepted by the passes,
    some slanted) are states the pipeline makes: the rule holds there too."""
    views = view_data(scene("corner320"))
    ov = oracle.OracleViews(views)
    ref = 2
    srcs = select_sources(scene("corner320").views, 4)[ref]
    opts = MvsOptions(num_sources=4, sweeps=N_PASSES // 2)
    planes = _init_planes(ov, ref, srcs, opts)
    mid, _ = ov.run_passes(ref, srcs, 1.0, 10.0, planes, 0, 3, N_PASSES, variant=0)
    assert (bits(mid[..., 0]) != 0).any()          # some planes are no longer fronto-parallel
    a, na = ov.run_passes(ref, srcs, 1.0, 10.0, mid, 3, N_PASSES, N_PASSES, variant=0)
    b, nb = ov.run_passes(ref, srcs, 1.0, 10.0, mid, 3, N_PASSES, N_PASSES, variant=5)
    assert np.array_equal(bits(a), bits(b))
    assert np.array_equal(na, nb)


def test_current_plane_rule_needs_the_planes_own_costs():
    """The constructed ring state stores cost 1 with a far plane whose real
    cost is lower.  The reference re-evaluates that plane when a neighbour
    repeats it and keeps it at its real cost; skipping it as a repeat of the
    current plane leaves the best at the planted 1, so a perturbation of the
    same real cost wins instead and the state changes -- which is why the
    device clears JobDesc.costs_exact for planted states."""
    ov = pins._scene()
    far = pins._fronto(pins.M + 5)
    p0 = pins._planes(far)
    p0[12, 24] = pins._fronto(pins.M, cost=0.0)
    a, _ = ov.run_passes(0, [1], pins.ZMIN, pins.ZMAX, p0, 0, 1, pins.N_PASSES, variant=0)
    b, _ = ov.run_passes(0, [1], pins.ZMIN, pins.ZMAX, p0, 0, 1, pins.N_PASSES, variant=5)
    differ = np.any(bits(a) != bits(b), axis=-1)
    assert differ.sum() > 100
    # where they differ the reference kept the far plane, re-costed below 1, and
    # the skipping run took another plane
    assert np.all(np.all(bits(a[differ][:, :4]) == bits(far[:4]), axis=-1))
    assert np.all(a[differ][:, 4] < 1.0)
    assert not np.any(np.all(bits(b[differ][:, :4]) == bits(far[:4]), axis=-1))
