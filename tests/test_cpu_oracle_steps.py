"""Known-answer tests of the oracle's half-sweep decision steps (FP64).

The pass tests compare the HIP path with the oracle; these pin the oracle
itself on hand-built inputs whose answers follow from the reference's text:
candidate search (lib 862-1104, src 720-913), joint view selection
(lib 1114-1183, src 915-981, transform_PDF_to_CDF src 79-92), final costs
(src 983-1002), adoption (lib 1278-1334, FindMinCostIndex /
FindMaxCostIndex lib 15-37), the refinement hypotheses and their sequential
adoption (lib 684-811) and the restricted score (lib 1268-1277, 784-793).
Each case is built so that a plausible misreading (a swapped 0.9/0.1 view
prior, a different threshold, first- instead of last-index ties, the wrong
flags, a different draw count, a probe off by one) changes an integer
outcome.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

FLT_EPSILON = 2.0 ** -23

# ------------------------------------------------------- candidate search ---

# Probe offsets (dy, dx) of each candidate for a pixel far from every border,
# read off the reference: far probes at 3 cells, then 5, 7, ..., 23
# ("up_far - 2 * i * width", i = 1..10); near probes at 1 cell, then the V
# pattern 2, 3, 4 cells out and 0, 1, 2 cells to either side.
INTERIOR = {
    0: [(-1, 0), (-2, 0), (-3, -1), (-3, 1), (-4, -2), (-4, 2)],
    1: [(-3 - 2 * i, 0) for i in range(11)],
    2: [(1, 0), (2, 0), (3, -1), (3, 1), (4, -2), (4, 2)],
    3: [(3 + 2 * i, 0) for i in range(11)],
    4: [(0, -1), (0, -2), (-1, -3), (1, -3), (-2, -4), (2, -4)],
    5: [(0, -3 - 2 * i) for i in range(11)],
    6: [(0, 1), (0, 2), (-1, 3), (1, 3), (-2, 4), (2, 4)],
    7: [(0, 3 + 2 * i) for i in range(11)],
}

# Pixels of a 7 x 7 grid near its borders, worked out by hand from the
# reference's bounds tests (p.y > 2 + 2i, p.x > i, p.x < width - 1 - i, ...).
# Missing keys: the candidate's flag is false.
BORDER_7x7 = {
    (1, 1): {  # (px, py)
        0: [(-1, 0)],
        2: [(1, 0), (2, 0), (3, 1), (4, 2)],
        3: [(3, 0), (5, 0)],
        4: [(0, -1)],
        6: [(0, 1), (0, 2), (1, 3), (2, 4)],
        7: [(0, 3), (0, 5)],
    },
    (5, 3): {
        0: [(-1, 0), (-2, 0), (-3, -1)],
        1: [(-3, 0)],
        2: [(1, 0), (2, 0), (3, -1)],
        3: [(3, 0)],
        4: [(0, -1), (0, -2), (-1, -3), (1, -3), (-2, -4), (2, -4)],
        5: [(0, -3), (0, -5)],
        6: [(0, 1)],
    },
    (0, 0): {
        2: [(1, 0), (2, 0), (3, 1), (4, 2)],
        3: [(3, 0), (5, 0)],
        6: [(0, 1), (0, 2), (1, 3), (2, 4)],
        7: [(0, 3), (0, 5)],
    },
    (6, 6): {
        0: [(-1, 0), (-2, 0), (-3, -1), (-4, -2)],
        1: [(-3, 0), (-5, 0)],
        4: [(0, -1), (0, -2), (-1, -3), (-2, -4)],
        5: [(0, -3), (0, -5)],
    },
    # the V probes never reach column 0 or column w - 1 (p.x > i,
    # p.x < width - 1 - i): up_near drops (-4, -2) here ...
    (2, 5): {
        0: [(-1, 0), (-2, 0), (-3, -1), (-3, 1), (-4, 2)],
        1: [(-3, 0), (-5, 0)],
        2: [(1, 0)],
        4: [(0, -1), (0, -2)],
        6: [(0, 1), (0, 2), (-1, 3), (-2, 4)],
        7: [(0, 3)],
    },
    # ... and (-4, 2) here
    (4, 5): {
        0: [(-1, 0), (-2, 0), (-3, -1), (-3, 1), (-4, -2)],
        1: [(-3, 0), (-5, 0)],
        2: [(1, 0)],
        4: [(0, -1), (0, -2), (-1, -3), (-2, -4)],
        5: [(0, -3)],
        6: [(0, 1), (0, 2)],
    },
}


def _probe_sets(h, w, px, py, compat=0):
    """The cells each candidate can pick, found by planting: with every cost
    1, a cell belongs to candidate k's probe set iff planting 0 there (2 for
    right_far, which keeps the maximum) makes k pick it."""
    sets = {k: set() for k in range(8)}
    flags = None
    for qy in range(h):
        for qx in range(w):
            if (qx, qy) == (px, py):
                continue
            for val, ks in ((0.0, (0, 1, 2, 3, 4, 5, 6)), (2.0, (7,))):
                g = np.ones((h, w))
                g[qy, qx] = val
                flag, pos = O.t_candidates(g, px, py, compat)
                flags = flag
                for k in ks:
                    if flag[k] and pos[k] == qy * w + qx:
                        sets[k].add((qy - py, qx - px))
    return flags, sets


def test_candidate_probe_sets_interior():
    h = w = 61
    flags, sets = _probe_sets(h, w, 30, 30)
    assert flags.all()
    for k in range(8):
        assert sets[k] == set(INTERIOR[k]), (k, sorted(sets[k]))


@pytest.mark.parametrize("pix", sorted(BORDER_7x7))
def test_candidate_probe_sets_7x7_borders(pix):
    px, py = pix
    flags, sets = _probe_sets(7, 7, px, py)
    want = BORDER_7x7[pix]
    assert [bool(f) for f in flags] == [k in want for k in range(8)]
    for k in range(8):
        assert sets[k] == set(want.get(k, [])), (k, sorted(sets[k]))


def test_candidate_defaults_and_ties_keep_the_nearest():
    """All probes equal: every candidate keeps its first probe. Equal minima:
    the nearer one is kept (strict <, lib 872). right_far keeps the first
    strict maximum (lib 954)."""
    h = w = 61
    px = py = 30
    c = py * w + px
    g = np.ones((h, w))
    flag, pos = O.t_candidates(g, px, py)
    assert list(pos) == [c - w, c - 3 * w, c + w, c + 3 * w, c - 1, c - 3, c + 1, c + 3]
    g[py - 5, px] = g[py - 9, px] = 0.5          # up_far: equal minima
    g[py + 3, px - 1] = g[py + 3, px + 1] = 0.5  # down_near: equal minima, left side tested first
    g[py, px + 5] = g[py, px + 9] = 2.0          # right_far: equal maxima
    g[py, px - 5] = 0.5                          # left_far: a deeper minimum ...
    g[py, px - 3] = 0.5                          # ... equal to the first probe
    flag, pos = O.t_candidates(g, px, py)
    assert pos[1] == c - 5 * w
    assert pos[2] == c + 3 * w - 1
    assert pos[7] == c + 5
    assert pos[5] == c - 3


def test_right_far_keeps_the_maximum_unless_fixed():
    h = w = 61
    px = py = 30
    c = py * w + px
    g = np.ones((h, w))
    g[py, px + 7] = 3.0
    g[py, px + 15] = 2.5
    g[py, px + 9] = 0.0
    g[py, px - 9] = 0.0  # left_far, for contrast: the minimum
    _, pos = O.t_candidates(g, px, py)
    assert pos[7] == c + 7 and pos[5] == c - 9
    _, pos = O.t_candidates(g, px, py, compat=1)  # GMVS_COMPAT_FIX_RIGHT_FAR
    assert pos[7] == c + 9 and pos[5] == c - 9


def test_far_probes_stop_at_23_cells():
    h = w = 61
    px = py = 30
    c = py * w + px
    g = np.ones((h, w))
    g[py - 23, px] = 0.3   # the deepest probe
    g[py - 25, px] = 0.05  # beyond reach
    g[py - 8, px] = 0.01   # even offset: never probed
    g[py - 9, px + 1] = 0.01  # off axis
    _, pos = O.t_candidates(g, px, py)
    assert pos[1] == c - 23 * w


# --------------------------------------------------------- view selection ---

def _table(cols):
    """[8, nv] candidate-major cost table from per-view columns."""
    return np.stack([np.asarray(col, np.float64) for col in cols], axis=1)


# Near flags up, down and left set; right and every far flag clear.
FLAGS_A = [True, False, True, False, True, False, False, False]
# Neighbour masks: up, down, left, right (right is ignored: its flag is clear).
NBSEL_A = [0b0001, 0b0011, 0b0101, 0b1111]


def _case_a():
    ca = _table([
        [0.0] * 8,                            # view 0: 8 good costs -> mean exp(0) = 1
        [1.5] * 8,                            # view 1: 8 costs > 1.2 -> 0
        [0.7] * 8,                            # view 2: 8 costs < tau = 0.8 -> exp(-0.49 / 0.18)
        [0.1, 0.1] + [1.0] * 6,               # view 3: only 2 below tau -> exp(-tau^2 / 0.32)
    ])
    # view priors, summed over the up, down and left neighbours (0.9 where the
    # neighbour selected the view, 0.1 where not)
    prior = np.array([0.9 * 3, 0.1 + 0.9 + 0.1, 0.1 + 0.1 + 0.9, 0.1 * 3])
    p = np.array([1.0, 0.0, math.exp(-0.49 / 0.18), math.exp(-0.64 / 0.32)]) * prior
    return ca, np.cumsum(p) / p.sum()


def test_view_selection_pdf_cdf_and_draws():
    ca, cdf_want = _case_a()
    # cdf_want = [0.95986, 0.95986, 0.98557, 1.0]
    assert cdf_want[0] < 0.97 < cdf_want[2] < 0.99
    draws = [0.5] * 6 + [0.9] * 3 + [0.97] * 3 + [0.99] * 2 + [1.0]
    cdf, wv, tsel, wn = O.t_select_views(ca, FLAGS_A, NBSEL_A, 0, draws)
    np.testing.assert_allclose(cdf, cdf_want, rtol=1e-12)
    # 0.5 and 0.9 fall in view 0, 0.97 in view 2, 0.99 and 1.0 in view 3 (a
    # draw of exactly 1 lands in the last view only because FLT_EPSILON is
    # subtracted from every draw)
    assert list(wv) == [9, 0, 3, 3]
    assert tsel == 0b1101
    assert wn == 15.0


def test_view_selection_draw_boundaries():
    """FLT_EPSILON is subtracted from every draw (lib 1161) and a view is
    taken where its CDF value is strictly above the result (lib 1166): a draw
    half an epsilon above a CDF step still lands below it, and a draw that
    lands exactly on the step goes to the next view."""
    ca, _ = _case_a()
    cdf, _, _, _ = O.t_select_views(ca, FLAGS_A, NBSEL_A, 0, [0.5] * 15)
    c0 = float(cdf[0])
    near = c0 + 0.5 * FLT_EPSILON
    on = c0 + FLT_EPSILON
    assert on - FLT_EPSILON == c0  # the oracle's subtraction lands exactly on the step
    _, wv, _, _ = O.t_select_views(ca, FLAGS_A, NBSEL_A, 0, [near] * 8 + [on] * 7)
    assert list(wv) == [8, 0, 7, 0]


def test_view_selection_threshold_decays_with_iteration():
    """tau = 0.8 exp(-it^2 / 90): a cost of 0.6 counts as good at it = 0
    (tau 0.8) and not at it = 6 (tau 0.536); with fewer than 3 good costs the
    view gets exp(-tau^2 / 0.32)."""
    ca = _table([[0.6] * 8, [0.0] * 8])
    flags = [True, False, True, False, True, False, True, False]
    nb = [0b11] * 4
    for it in (0, 6):
        tau = 0.8 * math.exp(-it * it / 90.0)
        p0 = math.exp(-0.36 / 0.18) if 0.6 < tau else math.exp(-tau * tau / 0.32)
        cdf, _, _, _ = O.t_select_views(ca, flags, nb, it, [0.5] * 15)
        np.testing.assert_allclose(cdf[0], p0 / (p0 + 1.0), rtol=1e-12)
    # draws between the two CDF values separate the iterations
    c0 = math.exp(-2.0) / (math.exp(-2.0) + 1)  # it = 0: 0.119
    tau6 = 0.8 * math.exp(-36 / 90.0)
    c6 = math.exp(-tau6 ** 2 / 0.32) / (math.exp(-tau6 ** 2 / 0.32) + 1)  # it = 6: 0.289
    mid = 0.5 * (c0 + c6)
    _, wv0, _, _ = O.t_select_views(ca, flags, nb, 0, [mid] * 15)
    _, wv6, _, _ = O.t_select_views(ca, flags, nb, 6, [mid] * 15)
    assert list(wv0) == [0, 15] and list(wv6) == [15, 0]


def test_view_selection_bad_count_and_good_count_boundaries():
    """A view with 3 costs > 1.2 gets probability 0 even with 5 good costs; a
    cost of exactly 1.2 is not bad; exactly 2 good costs fall back to
    exp(-tau^2 / 0.32), 3 good costs use their mean."""
    ca = _table([
        [0.1, 0.1, 0.1, 1.3, 1.3, 1.3, 0.5, 0.5],  # 3 bad -> 0
        [0.1, 0.1, 0.1, 1.3, 1.3, 1.2, 0.5, 0.5],  # 2 bad (1.2 is not > 1.2), 5 good
        [0.1, 0.1, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],  # 2 good -> exp(-2)
        [0.1, 0.1, 0.1, 1.0, 1.0, 1.0, 1.0, 1.0],  # 3 good -> exp(-0.01 / 0.18)
    ])
    flags = [True] * 8
    nb = [0b1111] * 4  # every view prior 3.6
    p = np.array([0.0,
                  (3 * math.exp(-0.01 / 0.18) + 2 * math.exp(-0.25 / 0.18)) / 5,
                  math.exp(-2.0),
                  math.exp(-0.01 / 0.18)])
    cdf, wv, tsel, wn = O.t_select_views(ca, flags, nb, 0, [1e-6] * 15)
    np.testing.assert_allclose(cdf, np.cumsum(p) / p.sum(), rtol=1e-12)
    assert cdf[0] == 0.0
    # tiny draws skip the zero-probability view 0
    assert list(wv) == [0, 15, 0, 0] and tsel == 0b10 and wn == 15


def test_view_selection_unflagged_neighbours_add_no_prior():
    """Only flagged near neighbours contribute a view prior; a view every
    counted neighbour selected outweighs one only an unflagged neighbour
    selected, even though its costs are worse."""
    ca = _table([[0.3] * 8, [0.0] * 8])
    flags = [False, True, True, True, False, True, False, True]  # only down_near of the near flags
    nb = [0b10, 0b01, 0b10, 0b10]  # the down neighbour selected view 0, the others view 1
    p = np.array([math.exp(-0.09 / 0.18) * 0.9, 1.0 * 0.1])
    cdf, wv, _, _ = O.t_select_views(ca, flags, nb, 0, [0.5] * 15)
    np.testing.assert_allclose(cdf, np.cumsum(p) / p.sum(), rtol=1e-12)
    assert list(wv) == [15, 0]


# ------------------------------------------------------ final costs, adoption ---

def test_final_costs_weight_only_drawn_views():
    ca = _table([
        [0.2, 0.4, 0.6, 0.8, 0.1, 0.3, 0.5, 0.7],
        [np.nan] * 8,  # not drawn: must not enter
        [0.9, 0.7, 0.5, 0.3, 0.2, 0.4, 0.6, 0.8],
    ])
    flags = [True, True, False, True, True, True, False, True]
    wv = [2.0, 0.0, 13.0]
    fc = O.t_final_costs(ca, flags, wv, 15.0)
    want = (2 * ca[:, 0] + 13 * ca[:, 2]) / 15
    np.testing.assert_allclose(fc, want, rtol=1e-14)
    # geometric term: flagged candidates add 0.1 * geom per drawn view,
    # unflagged ones add nothing (src 983-1002)
    geom = np.full((8, 3), 5.0)
    geom[:, 1] = np.nan
    fc = O.t_final_costs(ca, flags, wv, 15.0, geom)
    np.testing.assert_allclose(fc, want + np.where(flags, 0.5, 0.0), rtol=1e-14)


IN = [3.0] * 8  # candidate depths inside [1, 10]


def test_adopt_min_takes_the_last_of_equal_minima():
    fc = [0.5, 0.3, 0.3, 0.4, 0.9, 0.9, 0.9, 0.9]
    assert O.t_adopt(False, fc, [True] * 8, IN, 0.6, 1.0, 10.0) == 2


def test_adopt_min_checks_only_the_argmin():
    fc = [0.5, 0.3, 0.4, 0.9, 0.9, 0.9, 0.9, 0.9]
    dk = list(IN)
    dk[1] = 11.0  # the cheapest is out of range: nothing is adopted, not candidate 2
    assert O.t_adopt(False, fc, [True] * 8, dk, 0.6, 1.0, 10.0) == -1
    flags = [True] * 8
    flags[1] = False  # the cheapest is unflagged (its zero costs win the argmin)
    assert O.t_adopt(False, fc, flags, IN, 0.6, 1.0, 10.0) == -1


def test_adopt_min_strictly_cheaper_and_inclusive_range():
    fc = [0.5, 0.3, 0.4, 0.9, 0.9, 0.9, 0.9, 0.9]
    assert O.t_adopt(False, fc, [True] * 8, IN, 0.3, 1.0, 10.0) == -1  # equal: kept
    dk = list(IN)
    dk[1] = 1.0
    assert O.t_adopt(False, fc, [True] * 8, dk, 0.31, 1.0, 10.0) == 1
    dk[1] = 10.0
    assert O.t_adopt(False, fc, [True] * 8, dk, 0.31, 1.0, 10.0) == 1


def test_adopt_max_takes_the_last_of_equal_maxima():
    rs = [0.5, 0.8, 0.2, 0.8, 0.0, 0.0, 0.1, 0.0]
    assert O.t_adopt(True, rs, [True] * 8, IN, 0.6, 1.0, 10.0) == 3
    assert O.t_adopt(True, rs, [True] * 8, IN, 0.8, 1.0, 10.0) == -1  # not strictly higher
    dk = list(IN)
    dk[3] = 0.5
    assert O.t_adopt(True, rs, [True] * 8, dk, 0.6, 1.0, 10.0) == -1
    # unflagged candidates score 0: all zero -> the last index (7), unflagged
    assert O.t_adopt(True, [0.0] * 8, [True] * 7 + [False], IN, -1.0, 1.0, 10.0) == -1


# ------------------------------------------------------------- refinement ---

def test_refinement_hypotheses_order():
    nr, nc, npt = [0.1, 0.2, -0.97], [0.3, 0.0, -0.95], [0.0, 0.4, -0.9]
    d, n = O.t_refine_hypotheses(2.0, nr, 3.0, nc, 3.03, npt)
    assert list(d) == [2.0, 3.0, 2.0, 3.0, 3.03]
    assert np.array_equal(n, np.array([nc, nr, nr, npt, nc]))


def test_refinement_adoption_is_sequential():
    tc = [0.5, 0.4, 0.45, 0.3, 0.35]
    db = [3.0] * 5
    # each hypothesis is compared with the state the previous ones left
    assert O.t_refine_adopt(False, tc, db, [0] * 5, 1.0, 10.0, 0.6, 0.0)[:2] == (3, 0.3)
    db2 = list(db)
    db2[3] = 10.5  # out of range: skipped, and 0.35 then beats 0.4
    assert O.t_refine_adopt(False, tc, db2, [0] * 5, 1.0, 10.0, 0.6, 0.0)[:2] == (4, 0.35)
    assert O.t_refine_adopt(False, tc, db, [0] * 5, 1.0, 10.0, 0.3, 0.0)[0] == -1  # equal: kept


def test_refinement_prior_adoption_uses_the_restricted_score():
    tc = [0.9, 0.8, 0.7, 0.6, 0.5]
    sc = [0.2, 0.1, 0.5, 0.5, 0.3]
    db = [3.0] * 5
    # restricted score 0 (the propagation step adopted nothing): the first
    # in-range hypothesis is taken, then only strictly higher scores
    k, cost, r = O.t_refine_adopt(True, tc, db, sc, 1.0, 10.0, 0.4, 0.0)
    assert (k, cost, r) == (2, 0.7, 0.5)
    # a restricted score above every hypothesis keeps the state
    assert O.t_refine_adopt(True, tc, db, sc, 1.0, 10.0, 0.4, 0.6)[0] == -1


def test_prior_score_closed_forms():
    dmin, dmax = 1.0, 9.0
    sd = (dmax - dmin) / 64
    sa = math.pi * 5 / 180
    n = np.array([0.0, 0.6, -0.8])
    beta, gamma = 0.18, 0.5
    for c in (0.0, 0.3, 1.1):
        e = math.exp(-c * c / beta)
        # same plane as the prior
        assert O.t_prior_score(c, 4.0, 4.0, n, n, dmin, dmax) == pytest.approx(e * (gamma + 1), rel=1e-14)
        # one depth sigma away
        assert O.t_prior_score(c, 4.0 + sd, 4.0, n, n, dmin, dmax) == pytest.approx(
            e * (gamma + math.exp(-0.5)), rel=1e-12)
        # one angle sigma away (rotation by 5 degrees about x)
        rot = np.array([0.0, 0.6 * math.cos(sa) + 0.8 * math.sin(sa),
                        0.6 * math.sin(sa) - 0.8 * math.cos(sa)])
        assert O.t_prior_score(c, 4.0, 4.0, n, rot, dmin, dmax) == pytest.approx(
            e * (gamma + math.exp(-0.5)), rel=1e-9)
        # orthogonal normal: the prior term vanishes, gamma remains
        assert O.t_prior_score(c, 4.0, 4.0, n, np.array([1.0, 0, 0]), dmin, dmax) == pytest.approx(
            e * gamma, rel=1e-12)
    # a dot product rounded above 1 counts as angle 0 (clamped), not NaN
    m = np.array([0.6, 0.8, 0.0]) * (1 + 1e-15)
    assert O.t_prior_score(0.0, 4.0, 4.0, m, m, dmin, dmax) == pytest.approx(1.5, rel=1e-12)
