"""Host-side logic of the matching mirror (no GPU): the job-pair lists of
run_sfm (sfm-lite/src/lib.rs:197-212) and the output-slot layout of a batch."""
import numpy as np
import pytest

from rustcolmap_amd.matching import DEFAULT_RATIO, FrameMatcher, _desc, job_pairs


def test_job_pairs_sequential_window_of_three():
    # lib.rs:199-205: for i, for k in 1..=3, push (i, i + k) when in range
    assert job_pairs(5, True) == [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4)]
    assert job_pairs(1, True) == [] and job_pairs(0, True) == []


def test_job_pairs_exhaustive():
    # lib.rs:207-211: every i < j, i-major
    assert job_pairs(4, False) == [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    assert len(job_pairs(20, False)) == 190


def test_default_ratio_is_the_reference_default():
    assert DEFAULT_RATIO == 0.75   # Options::default (lib.rs:61)


def test_slots_layout():
    fm = FrameMatcher.__new__(FrameMatcher)   # slot arithmetic only, no device upload
    fm.counts = np.array([5, 0, 7, 3], np.int32)
    off = fm.slots([(0, 2), (1, 2), (2, 3), (3, 0), (0, 9)])
    # pair p owns min(n_i, n_j) slots; an out-of-range pair gets none (the library rejects it)
    np.testing.assert_array_equal(off, [0, 5, 5, 8, 11, 11])


def test_descriptor_shape_checks():
    assert _desc(np.zeros((0,), np.float32)).shape == (0, 128)
    assert _desc(np.ones((3, 128), np.float64)).dtype == np.float32
    with pytest.raises(ValueError):
        _desc(np.ones((3, 64), np.float32))
