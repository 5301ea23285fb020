"""The band split searches' region form of the split proposal.

The reference marks a pixel of the label for the new half when d1 > d2, its
4-connected distances inside the label from the two seeds, with d = -1 for a
seed that never reaches it (calc_seg_split, split_merge.cu:847-860).  The
band searches (bfs_bands in st_spix_amd/csrc/k_split.hip) instead grow two
regions level by level, each only into pixels neither holds yet, a pixel
reached by both at one level joining seed 1's region; the new half is seed
2's region when the two regions touch, else all of seed 1's.  This checks
that rule against the distance rule by brute force on random masks (several
pieces, seeds in one piece or in two, a missing second seed).
"""
from collections import deque

import numpy as np


def distances(M, s):
    d = np.full(M.shape, -1, dtype=np.int64)
    if s is None:
        return d
    H, W = M.shape
    d[s] = 0
    q = deque([s])
    while q:
        y, x = q.popleft()
        for yy, xx in ((y + 1, x), (y - 1, x), (y, x + 1), (y, x - 1)):
            if 0 <= yy < H and 0 <= xx < W and M[yy, xx] and d[yy, xx] < 0:
                d[yy, xx] = d[y, x] + 1
                q.append((yy, xx))
    return d


def grow(R):
    g = np.zeros_like(R)
    g[1:] |= R[:-1]
    g[:-1] |= R[1:]
    g[:, 1:] |= R[:, :-1]
    g[:, :-1] |= R[:, 1:]
    return g


def regions_rule(M, s1, s2):
    R1 = np.zeros_like(M)
    R2 = np.zeros_like(M)
    R1[s1] = True
    if s2 is not None:
        R2[s2] = True
    while True:
        free = M & ~(R1 | R2)
        N1 = grow(R1) & free
        N2 = grow(R2) & free & ~N1
        if not (N1.any() or N2.any()):
            break
        R1 |= N1
        R2 |= N2
    return R2 if (grow(R1) & R2).any() else R1


def test_region_rule_matches_distance_rule():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(800):
        H, W = (int(v) for v in rng.integers(2, 18, size=2))
        M = rng.random((H, W)) < rng.uniform(0.3, 0.95)
        pts = np.argwhere(M)
        if len(pts) < 2:
            continue
        s1 = tuple(pts[rng.integers(len(pts))])
        s2 = tuple(pts[rng.integers(len(pts))]) if rng.random() < 0.9 else None
        if s2 == s1:
            continue
        d1, d2 = distances(M, s1), distances(M, s2)
        want = (d1 >= 0) & (d1 > d2)
        assert np.array_equal(regions_rule(M, s1, s2), want)
        checked += 1
    assert checked > 600


def test_region_rule_ties_stay_with_seed_one():
    # a 1 x 5 strip with the seeds at both ends: the middle pixel is a tie
    M = np.ones((1, 5), dtype=bool)
    K = regions_rule(M, (0, 0), (0, 4))
    assert K.tolist() == [[False, False, False, True, True]]
    # seed 2 in another piece: all of seed 1's piece moves
    M = np.array([[1, 1, 0, 1, 1]], dtype=bool)
    K = regions_rule(M, (0, 0), (0, 4))
    assert K.tolist() == [[True, True, False, False, False]]
