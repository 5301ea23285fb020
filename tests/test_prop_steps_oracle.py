"""Pins the oracle's restatements of the prop_cuda per-step ops
(shift_labels.cu:28-62, fill_missing.cu:49-314) with hand-worked cases
(translation, collisions, the tie rule, the checkerboard order, the stop
rule that compares the first round with 1) and an independent numpy
scatter formulation of shift_labels."""
import numpy as np


def shift_ref(spix, flow):
    """numpy scatter form: arrivals by np.add.at, largest label by
    np.maximum.at (not the oracle's pixel loop)."""
    H, W = spix.shape
    yy, xx = np.mgrid[0:H, 0:W]
    L = spix.ravel()
    ok = (L >= 0) & (L < flow.shape[0])
    hd = yy.ravel()[ok] + flow[L[ok], 1]
    wd = xx.ravel()[ok] + flow[L[ok], 0]
    inb = (hd >= 0) & (hd < H) & (wd >= 0) & (wd < W)
    d = hd[inb] * W + wd[inb]
    counts = np.zeros(H * W, np.int64)
    np.add.at(counts, d, 1)
    shifted = np.full(H * W, -1, np.int64)
    np.maximum.at(shifted, d, L[ok][inb])
    return shifted.reshape(H, W), counts.reshape(H, W)


def test_shift_translation_and_collisions(orc):
    # one label translated by (+2, -1): a shifted copy, counts 1 inside
    spix = np.zeros((5, 7), np.int32)
    sh, ct = orc.prop_shift_labels(spix, np.array([[2, -1]], np.int32))
    exp = np.full((5, 7), -1, np.int32)
    exp[0:4, 2:7] = 0
    np.testing.assert_array_equal(sh, exp)
    np.testing.assert_array_equal(ct, (exp == 0).astype(np.int32))
    # two labels landing on the same column: counts 2, the larger label wins
    spix = np.array([[0, 1, 2]], np.int32)
    flow = np.array([[1, 0], [0, 0], [-1, 0]], np.int32)
    sh, ct = orc.prop_shift_labels(spix, flow)
    np.testing.assert_array_equal(sh, [[-1, 2, -1]])
    np.testing.assert_array_equal(ct, [[0, 3, 0]])
    # a label without a flow entry stays where it is not (it does not move)
    spix = np.array([[0, 5]], np.int32)
    sh, ct = orc.prop_shift_labels(spix, np.array([[1, 0]], np.int32))
    np.testing.assert_array_equal(sh, [[-1, 0]])
    np.testing.assert_array_equal(ct, [[0, 1]])


def test_shift_random_against_scatter(orc):
    rng = np.random.default_rng(3)
    for it in range(20):
        H, W = rng.integers(1, 40, 2)
        n = int(rng.integers(1, 30))
        spix = rng.integers(-1, n + 2, (H, W)).astype(np.int32)
        flow = rng.integers(-6, 7, (n, 2)).astype(np.int32)
        sh, ct = orc.prop_shift_labels(spix, flow)
        rs, rc = shift_ref(spix, flow)
        np.testing.assert_array_equal(sh, rs)
        np.testing.assert_array_equal(ct, rc)


def holes(seg):
    return np.flatnonzero(seg.ravel() < 0).astype(np.int32)


def test_fill_row_worked_by_hand(orc):
    # [0 -1 -1 -1 1], centres at x = 0 and x = 4.  Round 0: pixels 1 and 3
    # border (phase x odd) -> 0 and 1.  Round 1: pixel 2, equal distance to
    # both centres; strict '>' keeps the first candidate in N, S, W, E
    # order, W = 0.  Round 2 counts 0 holes and stops: 3 rounds.
    seg = np.array([[0, -1, -1, -1, 1]], np.int32)
    c = np.array([[0, 0], [4, 0]], np.float32)
    out, rounds, neg = orc.prop_fill_missing(seg, c, holes(seg), 2000)
    np.testing.assert_array_equal(out, [[0, 0, 0, 1, 1]])
    assert (rounds, neg) == (3, 0)


def test_fill_checkerboard_order(orc):
    # a 2x2 hole under label 0 with label 1 to the right; W = 4, H = 3,
    # centres 0: (1.5, 0), 1: (3, 1.5).  All four holes are border at round
    # 0; the phases (x%2, y%2) run (0,0), (0,1), (1,0), (1,1) and each sees
    # what the earlier phases of the round wrote:
    #   (2,2) phase (0,0): N, W holes, E = 1                    -> 1
    #   (2,1) phase (0,1): N = 0: -(0.125^2 + 0.3333^2) = -0.1267,
    #                      S = E = 1: -(0.25^2 + 0.1667^2) = -0.0903  -> 1
    #   (1,2) phase (1,0): W = 0: -(0.125^2 + 0.6667^2) = -0.4601,
    #                      E = 1 (from phase (0,0)): -(0.5^2 + 0.1667^2) = -0.2778 -> 1
    #   (1,1) phase (1,1): N = W = 0: -0.1267, S = E = 1: -0.2778   -> 0
    # round 1 counts no hole: 2 rounds
    seg = np.array([[0, 0, 0, 0],
                    [0, -1, -1, 1],
                    [0, -1, -1, 1]], np.int32)
    c = np.array([[1.5, 0.0], [3.0, 1.5]], np.float32)
    out, rounds, neg = orc.prop_fill_missing(seg, c, holes(seg), 2000)
    np.testing.assert_array_equal(out, [[0, 0, 0, 0], [0, 0, 1, 1], [0, 1, 1, 1]])
    assert neg == 0 and rounds == 2


def test_fill_stop_rules(orc):
    z = np.zeros((1, 2), np.float32)
    # one hole: the first round counts 1, equal to the initial comparison
    # value 1, so the loop ends after that round, which has filled it
    seg = np.array([[0, -1]], np.int32)
    out, rounds, neg = orc.prop_fill_missing(seg, z, holes(seg), 2000)
    np.testing.assert_array_equal(out, [[0, 0]])
    assert (rounds, neg) == (1, 0)
    # no labelled pixel: rounds count 9, then 9 again (no shrink) -> 2 rounds
    seg = np.full((3, 3), -1, np.int32)
    out, rounds, neg = orc.prop_fill_missing(seg, z, holes(seg), 2000)
    assert (rounds, neg) == (2, 9)
    np.testing.assert_array_equal(out, seg)
    # break_iter caps the rounds: a long hole row fills from one end
    seg = np.full((1, 9), -1, np.int32)
    seg[0, 0] = 0
    out, rounds, neg = orc.prop_fill_missing(seg, z, holes(seg), 3)
    assert rounds == 3
    # rounds fill 1, 1 (phase x even after odd), ...: x = 1 in round 0, x = 2
    # in round 1, x = 3 in round 2 (one pixel per round along a row)
    np.testing.assert_array_equal(out[0, :4], [0, 0, 0, 0])
    assert (out[0, 4:] == -1).all() and neg == 5


def test_fill_random_invariants(orc):
    """Random label maps with holes: when every hole component touches a
    label, nothing is left; each filled pixel carries a label that one of
    its 4-neighbours has."""
    rng = np.random.default_rng(5)
    for it in range(10):
        H, W = rng.integers(8, 40, 2)
        lab = (np.arange(H)[:, None] // 5 * 10 + np.arange(W)[None, :] // 7).astype(np.int32)
        seg = lab.copy()
        seg[rng.random((H, W)) < 0.4] = -1
        seg[0, :] = lab[0, :]  # a labelled row keeps every hole reachable
        n = int(lab.max()) + 1
        cen = rng.uniform(0, max(H, W), (n, 2)).astype(np.float32)
        out, rounds, neg = orc.prop_fill_missing(seg, cen, holes(seg), 2000)
        assert neg == 0 and (out >= 0).all()
        assert (out[seg >= 0] == seg[seg >= 0]).all()
        pad = np.pad(out, 1, constant_values=-2)
        nb = np.stack([pad[:-2, 1:-1], pad[2:, 1:-1], pad[1:-1, :-2], pad[1:-1, 2:]])
        filled = seg < 0
        assert ((nb == out[None]).any(0) | ~filled).all()
