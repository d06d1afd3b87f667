"""Model of the matching tile kernel's cross-lane two-smallest merges
(rustcolmap_amd/csrc/match.hip: two_min_tstep / two_min_rows16 for the
rows, two_min_xlane for the columns), checked against a direct two-smallest
over all lanes.

Each of the 16 lanes of a row group holds a (u1 <= u2) pair for each of its
16 rows.  The kernel's tree: at each step a lane keeps half of its rows,
sends the other half to its partner under a DPP lane pattern and merges
what the partner sends.  The patterns are row_mirror (lane r <- 15 - r),
row_half_mirror (r <- (r & 8) | (7 - (r & 7))), quad_perm [2,3,0,1]
(r <- r ^ 2) and quad_perm [1,0,3,2] (r <- r ^ 1).  Lane r must end with row
r merged over all 16 lanes.  The column merge swaps a register with a copy
of itself across 16- or 32-lane halves (permlane16/32_swap), which leaves
each lane its own value and its partner's; the merge is symmetric.
"""
import numpy as np
import pytest


def two_min_merge(a1, a2, b1, b2):
    return np.minimum(a1, b1), np.minimum(np.maximum(a1, b1), np.minimum(a2, b2))


def direct_two_min(values):
    """(smallest, second smallest) of a multiset of values, per row."""
    s = np.sort(values, axis=-1)
    return s[..., 0], s[..., 1]


PATTERNS = {
    "row_mirror": lambda r: 15 - r,
    "row_half_mirror": lambda r: (r & 8) | (7 - (r & 7)),
    "quad_2301": lambda r: r ^ 2,
    "quad_1032": lambda r: r ^ 1,
}
KERNEL_STEPS = [("row_mirror", 8, 8), ("row_half_mirror", 4, 4), ("quad_2301", 2, 2), ("quad_1032", 1, 1)]


def tree(a, b, steps=KERNEL_STEPS, swap_halves=False):
    """a, b: (16 lanes, 16 rows) pairs; returns each lane's pair in slot 0.
    steps: (pattern, H, hi bit) as two_min_rows16 issues them."""
    a = a.copy()
    b = b.copy()
    lanes = np.arange(16)
    for pat, H, bit in steps:
        src = np.array([PATTERNS[pat](r) for r in lanes])
        hi = (lanes & bit) != 0
        if swap_halves:
            hi = ~hi
        na, nb = a.copy(), b.copy()
        for t in range(H):
            ka = np.where(hi, a[:, t + H], a[:, t])
            kb = np.where(hi, b[:, t + H], b[:, t])
            sa = np.where(hi, a[:, t], a[:, t + H])
            sb = np.where(hi, b[:, t], b[:, t + H])
            pa, pb = sa[src], sb[src]   # DPP: lane r reads the partner's register
            na[:, t], nb[:, t] = two_min_merge(ka, kb, pa, pb)
        a, b = na, nb
    return a[:, 0], b[:, 0]


def lane_pairs(rng, shape, ties=False):
    """Per-lane (u1 <= u2) pairs of 4 values each (the kernel's per-lane
    two-smallest over its 4 column blocks), with +inf for padded columns."""
    v = rng.integers(0, 6 if ties else 1 << 20, shape + (4,)).astype(np.float32)
    v[rng.random(shape + (4,)) < 0.1] = np.inf
    s = np.sort(v, axis=-1)
    return s[..., 0], s[..., 1], v


@pytest.mark.parametrize("seed,ties", [(0, False), (1, True), (2, True), (3, False)])
def test_row_tree_ends_with_own_row_merged_over_16_lanes(seed, ties):
    rng = np.random.default_rng(seed)
    a, b, v = lane_pairs(rng, (16, 16), ties)          # [lane, row]
    ga, gb = tree(a, b)
    # row r over all 16 lanes' 4 values each
    want1, want2 = direct_two_min(v.transpose(1, 0, 2).reshape(16, 64))
    np.testing.assert_array_equal(ga, want1)
    np.testing.assert_array_equal(gb, want2)


def test_row_tree_partners_hold_the_same_rows():
    """At every step the partner has the opposite hi bit and the same earlier
    ones, so it keeps the rows this lane sends."""
    lanes = np.arange(16)
    for k, (pat, H, bit) in enumerate(KERNEL_STEPS):
        p = np.array([PATTERNS[pat](r) for r in lanes])
        assert np.all((p & bit) != (lanes & bit))
        for _, _, earlier in KERNEL_STEPS[:k]:
            assert np.all((p & earlier) == (lanes & earlier))
        assert np.all(p[p] == lanes)   # a pairing


@pytest.mark.parametrize("mistake", ["swapped_halves", "dropped_step", "xor8_for_mirror_order"])
def test_row_tree_planted_mistakes_fail(mistake):
    rng = np.random.default_rng(5)
    a, b, v = lane_pairs(rng, (16, 16))
    want1, want2 = direct_two_min(v.transpose(1, 0, 2).reshape(16, 64))
    if mistake == "swapped_halves":
        ga, gb = tree(a, b, swap_halves=True)
    elif mistake == "dropped_step":
        ga, gb = tree(a, b, steps=KERNEL_STEPS[:3])
    else:
        # the quad steps first: partners then differ in low bits while the
        # kept rows are chosen by the high ones
        ga, gb = tree(a, b, steps=[("quad_1032", 8, 8), ("quad_2301", 4, 4), ("row_half_mirror", 2, 2),
                                   ("row_mirror", 1, 1)])
    assert not (np.array_equal(ga, want1) and np.array_equal(gb, want2))


def permlane_swap(x, span):
    """v_permlane16/32_swap of a register with a copy of itself: returns the
    two registers afterwards (64 lanes)."""
    v0, v1 = x.copy(), x.copy()
    if span == 32:
        v0[32:], v1[:32] = x[:32], x[32:]
    else:
        for row in (1, 3):              # odd rows of the first operand
            v0[16 * row:16 * row + 16] = x[16 * (row - 1):16 * row]
        for row in (0, 2):              # even rows of the second operand
            v1[16 * row:16 * row + 16] = x[16 * (row + 1):16 * (row + 2)]
    return v0, v1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_column_merges_cover_the_four_lane_groups(seed):
    """two_min_xlane<16> then <32>: lane l ends with the merge over lanes
    l, l ^ 16, l ^ 32, l ^ 48 (the 4 kq groups of a column)."""
    rng = np.random.default_rng(seed)
    u1, u2, v = lane_pairs(rng, (64,), ties=seed == 1)
    for span in (16, 32):
        s1, s2 = permlane_swap(u1, span), permlane_swap(u2, span)
        # each lane holds its own value and the partner's, in some order
        lanes = np.arange(64)
        assert np.all((s1[0] == u1) | (s1[1] == u1))
        assert np.all(np.sort(np.stack(s1), 0) == np.sort(np.stack([u1, u1[lanes ^ span]]), 0))
        u1, u2 = two_min_merge(s1[0], s2[0], s1[1], s2[1])
    col = np.arange(64) & 15
    for c in range(16):
        want1, want2 = direct_two_min(v[col == c].reshape(-1))
        assert np.all(u1[col == c] == want1) and np.all(u2[col == c] == want2)
