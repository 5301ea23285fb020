"""Pins single steps of the oracle against independent restatements of the
reference (numpy / scipy, written from the reference's source, not from the
oracle), brute force and library routines.  Inputs are chosen so that a
plausible slip (a dropped term, a wrong sign, a transposed index, a wrong
tie rule) changes the result."""
import math
from collections import deque

import numpy as np
import pytest
from scipy.special import gammaln

from st_spix_amd import synth

f32 = np.float32
FX = 2.0 ** 32


# ---------------------------------------------------------------- calc_joint
def joint_ref(img, x, y, mu, mshape, isig, logdet, count, s2, neq, potts):
    """update_seg.cu:194-259 in float32 operation order (no contraction)."""
    res = f32(-1000.0)
    x0, x1, x2 = (f32(img[c]) - f32(mu[c]) for c in range(3))
    d0 = int(math.trunc(float(x) - mshape[0]))  # int = int - double, truncated
    d1 = int(math.trunc(float(y) - mshape[1]))
    sx, sy, sz = (f32(v) for v in isig)
    res = f32(res - f32(f32(f32(x0 * x0) + f32(x1 * x1)) + f32(x2 * x2)) / f32(s2))
    res = f32(res - f32(f32(d0 * d0) * sx))
    res = f32(res - f32(f32(d1 * d1) * sz))
    res = f32(res - f32(f32(2 * d0 * d1) * sy))
    res = f32(res - f32(logdet))
    res = f32(res - f32(f32(potts) * f32(neq)))
    res = f32(res - f32(100 * (count <= 10)))
    return float(res)


def test_calc_joint_every_term(orc):
    rng = np.random.default_rng(1)
    for it in range(400):
        img = rng.uniform(0, 1, 3).astype(np.float32)
        mu = rng.uniform(0, 1, 3).astype(np.float32)
        mshape = rng.uniform(-5, 60, 2)
        x, y = int(rng.integers(0, 64)), int(rng.integers(0, 64))
        a, c = rng.uniform(0.01, 0.5, 2)
        b = rng.uniform(-0.8, 0.8) * math.sqrt(a * c)
        isig = np.array([a, b, c])
        logdet = rng.uniform(-3, 6)
        count = int(rng.choice([1, 5, 10, 11, 400]))
        neq = float(rng.integers(0, 9))
        potts = float(rng.choice([0.0, 0.5, 10.0]))
        got = orc.calc_joint(img, x, y, mu, mshape, isig, logdet, count, 8.1e-5, neq, potts)
        ref = joint_ref(img, x, y, mu, mshape, isig, logdet, count, 8.1e-5, neq, potts)
        assert got == ref, (it, got, ref)
    # each term matters on these inputs: the cross term's sign and the
    # int truncation toward zero (x - mu < 0) change the score
    img, mu = np.array([0.3, 0.2, 0.1], np.float32), np.array([0.25, 0.2, 0.1], np.float32)
    base = dict(img=img, x=3, y=9, mu=mu, mshape=np.array([5.7, 4.2]),
                isig=np.array([0.1, 0.05, 0.2]), logdet=2.5, count=8, s2=8.1e-5, neq=3.0,
                potts=10.0)
    ref = joint_ref(**base)
    assert orc.calc_joint(*base.values()) == ref
    flipped = dict(base, isig=np.array([0.1, -0.05, 0.2]))
    assert joint_ref(**flipped) != ref
    floor_d0 = math.floor(3 - 5.7)  # -3, the reference truncates to -2
    assert floor_d0 != int(math.trunc(3 - 5.7))


# ---------------------------------------------------------------- NIG marginal
def nig_consts(s2):
    a0 = f32(1e4)
    b0 = f32(f32(np.sqrt(f32(s2))) * f32(2.0)) * a0
    return float(a0), float(f32(b0))


def nig_ref(s2, n, bn):
    """calc_split_stats_step1_p (split_merge_prop.cu:2934-3021) and the
    channel sum of update_split_flag_p (:3052-3070) in FP64, the lgamma terms
    kept: a_0 is overwritten by a_n before both lgamma calls."""
    a0, b0 = nig_consts(s2)
    a_n = a0 + n / 2.0
    v_n = 1.0 / n
    num = a_n * math.log(b0) + gammaln(a_n) + 0.5 * math.log(v_n)
    tot = 0.0
    for c in range(3):
        den = a_n * math.log(bn[c]) + 0.5 * n * math.log(math.pi) + n * math.log(2) + gammaln(a_n)
        tot += num - den
    return tot


@pytest.mark.parametrize("s2", [8.1e-5, 0.01, 0.1])
def test_nig_marginal_against_gammaln(orc, s2):
    rng = np.random.default_rng(2)
    for n in (1, 2, 7, 100, 2500, 250000):
        for _ in range(20):
            bn = rng.uniform(0.1, 5e4, 3)
            got, ref = orc.nig_marginal(s2, n, bn), nig_ref(s2, n, bn)
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (n, got, ref)


# ---------------------------------------------------------------- split
def bfs(mask, seed):
    d = -np.ones(mask.shape, np.int64)
    if seed is None:
        return d
    q = deque([seed])
    d[seed] = 1
    H, W = mask.shape
    while q:
        y, x = q.popleft()
        for yy, xx in ((y - 1, x), (y, x - 1), (y + 1, x), (y, x + 1)):
            if 0 <= yy < H and 0 <= xx < W and mask[yy, xx] and d[yy, xx] < 0:
                d[yy, xx] = d[y, x] + 1
                q.append((yy, xx))
    return d


def seed_of(seg, k, mx, my, direction):
    """init_split(_p) (split_merge.cu:803-845): the seed's linear index is
    checked, not its column, so x = -1 reaches the previous row."""
    H, W = seg.shape
    if direction in (1, -1):
        x, y = int(mx) + direction, int(my)
    else:
        x, y = int(mx), int(my) + direction
    ind = y * W + x
    if ind < 0 or ind > W * H - 1 or seg.flat[ind] != k:
        return None
    return (ind // W, ind % W)


def fx_sums(v):
    """Per-pixel 2^-32 fixed-point sums (the oracle's statistics rule)."""
    v = v.astype(np.float64)
    return (np.rint(v * FX).sum() / FX, np.rint(v * v * FX).sum() / FX)


def split_ref(img, seg, K, direction, s2, split_alpha=0.0, prop=None, shifted=None):
    H, W = seg.shape
    yy, xx = np.mgrid[0:H, 0:W]
    halves = seg.copy()
    hasting = np.full(K, -999999.0)
    accept = np.zeros(K, np.int32)
    a0, b0 = nig_consts(s2)
    for k in range(K):
        m = seg == k
        n = int(m.sum())
        if n == 0:
            continue
        mx, my = xx[m].sum() / n, yy[m].sum() / n
        d1 = bfs(m, seed_of(seg, k, mx, my, direction))
        d2 = bfs(m, seed_of(seg, k, mx, my, -direction))
        halves[m & (d1 > d2)] = k + K  # calc_seg_split, split_merge.cu:847-860
        mk, ms = m & (d1 <= d2), m & (d1 > d2)
        nk, ns = int(mk.sum()), int(ms.sum())
        if nk < 1 or ns < 1:
            continue
        marg = []
        for part, cnt, full in ((mk, nk, False), (ms, ns, False), (m, n, True)):
            bn = []
            for c in range(3):
                if full:  # b_n_f: squares of both halves, sum of the whole label
                    qk = fx_sums(img[mk][:, c])[1] + fx_sums(img[ms][:, c])[1]
                    s = fx_sums(img[m][:, c])[0]
                    bn.append(b0 + 0.5 * (qk - s * s / cnt))
                else:
                    s, q = fx_sums(img[part][:, c])
                    bn.append(b0 + 0.5 * (q - s * s / cnt))
            marg.append(nig_ref(s2, cnt, bn))
        h = marg[0] + marg[1] - marg[2] + split_alpha
        if prop is not None:  # update_split_flag_p, split_merge_prop.cu:3053-3068
            inv_k = int(((shifted < 0) & mk).sum())
            inv_s = int(((shifted < 0) & ms).sum())
            propagate_k = inv_k <= inv_s
            iperc = inv_k / nk if propagate_k else inv_k / ns  # ninvalid_k in both
            iperc = iperc if prop[k] else 1.0
            h = h - 2.0 * (1.0 - iperc)
        hasting[k] = h
        accept[k] = h > -2.0
    return halves, hasting, accept


def voronoi_labels(H, W, K, seed):
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(0, W, K), rng.uniform(0, H, K)], -1)
    yy, xx = np.mgrid[0:H, 0:W]
    d = (xx[..., None] - pts[:, 0]) ** 2 + (yy[..., None] - pts[:, 1]) ** 2
    return np.argmin(d, -1).astype(np.int32)


@pytest.mark.parametrize("direction", [1, 2])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_split_proposal_and_hastings(orc, direction, seed):
    H, W, K = 36, 44, 9
    seg = voronoi_labels(H, W, K, seed)
    if seed == 2:  # a label in two pieces: the far piece is never reached
        seg[0:5, 0:5] = seg[H - 1, W - 1]
    rng = np.random.default_rng(10 + seed)
    img = (rng.uniform(0, 1, (K, 3))[seg] + rng.normal(0, 0.03, (H, W, 3))).astype(np.float32)
    img[seg == 3] += np.float32(0.2) * (np.mgrid[0:H, 0:W][1][seg == 3, None] > W // 2)
    halves, hasting, accept, _ = orc.split_eval(img, seg, K, direction, 8.1e-5)
    h_r, hs_r, acc_r = split_ref(img, seg, K, direction, 8.1e-5)
    np.testing.assert_array_equal(halves, h_r)
    ok = hs_r > -999999.0
    np.testing.assert_allclose(hasting[ok], hs_r[ok], rtol=1e-11, atol=1e-9)
    np.testing.assert_array_equal(accept, acc_r)
    assert (~ok).sum() < K  # most labels evaluate


def test_split_hastings_prop_mode(orc):
    H, W, K = 40, 40, 7
    seg = voronoi_labels(H, W, K, 5)
    rng = np.random.default_rng(6)
    img = (rng.uniform(0, 1, (K, 3))[seg] + rng.normal(0, 0.03, (H, W, 3))).astype(np.float32)
    shifted = seg.copy()
    shifted[rng.uniform(0, 1, (H, W)) < 0.15] = -1
    prop = (np.arange(K) % 2).astype(np.int32)
    halves, hasting, accept, _ = orc.split_eval(img, seg, K, 1, 8.1e-5, prop=prop,
                                                shifted=shifted, prop_mode=1)
    h_r, hs_r, acc_r = split_ref(img, seg, K, 1, 8.1e-5, prop=prop, shifted=shifted)
    np.testing.assert_array_equal(halves, h_r)
    ok = hs_r > -999999.0
    np.testing.assert_allclose(hasting[ok], hs_r[ok], rtol=1e-11, atol=1e-9)
    np.testing.assert_array_equal(accept, acc_r)


# ---------------------------------------------------------------- merge
def merge_ref(img, seg, K, direction, s2, alpha, merge_alpha=0.0, prop=None):
    """calc_merge_candidate(_p), calc_merge_stats_step0/1/2(_p) and
    update_merge_flag(_p) (split_merge_prop.cu:1808-1845, 2572-2766) in FP64;
    the winner of a target is the largest accepting k."""
    H, W = seg.shape
    prop_mode = prop is not None
    cand = -np.ones(K, np.int64)
    for y in range(H):
        for x in range(W):
            C = seg[y, x]
            inner = 1 <= y < H - 1 and 1 <= x < W - 1
            if inner and C == seg[y - 1, x] == seg[y + 1, x] == seg[y, x - 1] == seg[y, x + 1]:
                continue  # not a border pixel
            Wn = -1
            if direction == 1:
                if (1 < y < H - 2) if prop_mode else (y < H - 1):
                    Wn = seg[y + 1, x]
            else:
                if (1 < x < W - 2) if prop_mode else (x > 0):
                    Wn = seg[y, x - 1]
            if Wn >= 0 and Wn != C and not (prop_mode and prop[C] == prop[Wn]):
                cand[C] = max(cand[C], Wn)
    a0, b0 = nig_consts(s2)
    log_alpha = math.log(float(f32(math.exp(alpha))))
    sums = [[fx_sums(img[seg == k][:, c]) for c in range(3)] for k in range(K)]
    cnt = np.bincount(seg.ravel(), minlength=K)
    own = {}
    for k in range(K):
        if cand[k] < 0:
            continue
        own[k] = [max(b0 + 0.5 * (sums[k][c][1] - sums[k][c][0] ** 2 / cnt[k]), 0.1)
                  if b0 + 0.5 * (sums[k][c][1] - sums[k][c][0] ** 2 / cnt[k]) >= 0 else 0.1
                  for c in range(3)]
    hasting = np.full(K, -999999.0)
    winner = -np.ones(K, np.int64)
    for k in range(K):
        f = cand[k]
        if f < 0:
            continue
        n = cnt[f] + cnt[k]
        bf = []
        for c in range(3):
            v = b0 + 0.5 * ((sums[k][c][1] + sums[f][c][1]) -
                            (sums[f][c][0] + sums[k][c][0]) ** 2 / n)
            bf.append(v if v >= 0 else 0.1)
        if f not in own:  # the target never ran step0: b_n = 0, marginal +inf
            hasting[k] = -math.inf
            continue
        h = nig_ref(s2, n, bf) - (log_alpha + nig_ref(s2, cnt[k], own[k]) +
                                  nig_ref(s2, cnt[f], own[f])) + merge_alpha
        hasting[k] = h
        if h > -2.0:
            winner[f] = max(winner[f], k)
    return cand, hasting, winner


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("prop_mode", [0, 1])
def test_merge_candidates_and_hastings(orc, direction, prop_mode):
    H, W, K = 30, 38, 10
    seg = voronoi_labels(H, W, K, 23)  # geometry with accepted and rejected merges
    rng = np.random.default_rng(8)
    base = rng.uniform(0, 1, (K, 3))
    base[:8] = base[0]  # eight of ten labels share a colour: adjacent pairs merge
    img = (base[seg] + rng.normal(0, 0.01, (H, W, 3))).astype(np.float32)
    prop = (np.arange(K) % 2).astype(np.int32) if prop_mode else None
    cand, hasting, winner = orc.merge_eval(img, seg, K, direction, 8.1e-5, math.log(0.5),
                                           prop=prop, prop_mode=prop_mode)
    c_r, h_r, w_r = merge_ref(img, seg, K, direction, 8.1e-5, math.log(0.5), prop=prop)
    np.testing.assert_array_equal(cand, c_r)
    fin = np.isfinite(h_r) & (h_r > -999999.0)
    np.testing.assert_allclose(hasting[fin], h_r[fin], rtol=1e-11, atol=1e-6)
    assert (hasting[np.isneginf(h_r)] < -1e300).all()
    np.testing.assert_array_equal(winner, w_r)
    assert (w_r >= 0).any()  # some merge is accepted
    assert (h_r[fin] <= -2.0).any()  # and some rejected


# ---------------------------------------------------------------- relabel
def test_relabel_search_bruteforce(orc):
    """find_most_similar_spix + mark_for_relabel (relabel.cu:125-276): each
    living label takes the nearest dead previous id (ties to the smaller
    id) if closer than thresh_relabel, else a new id when its own previous
    colour moved by more than thresh_new; new ids in ascending label order."""
    rng = np.random.default_rng(4)
    nprev, nspix, nbuf = 30, 45, 200
    prior = rng.uniform(0, 1, (nprev, 3)).astype(np.float32)
    mu = rng.uniform(0, 1, (nspix, 3)).astype(np.float32)
    living = (rng.uniform(0, 1, nspix) < 0.7).astype(np.uint8)
    dead = [c for c in range(nprev) if not living[c]]
    for s in range(nspix):  # near copies of dead and own colours
        r = rng.uniform()
        if r < 0.3 and dead:
            mu[s] = prior[dead[int(rng.integers(0, len(dead)))]] + f32(1e-4)
        elif r < 0.5 and s < nprev:
            mu[s] = prior[s] + f32(1e-3)
    mu[44] = prior[dead[0]]  # two living labels tied on one dead id
    mu[43] = prior[dead[0]]
    living[[43, 44]] = 1
    target, nxt = orc.relabel_search(mu, living, prior, nprev, 1e-5, 1e-2, nbuf)
    exp = -np.ones(nspix, np.int64)
    nx = nspix - 1
    for s in range(nspix):
        if not living[s]:
            continue
        best, cand, ss = f32(10000.0), -1, f32(0.0)
        for c in list(range(nprev)):
            if living[c] and c != s:
                continue
            d = prior[c] - mu[s]
            delta = f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))
            if c == s:
                ss = delta
            if delta < best:
                best, cand = delta, c
        if cand >= 0 and best < f32(1e-5) and cand < nspix and cand != s:
            exp[s] = cand
        elif ss > f32(1e-2) and nx + 1 < nbuf:
            nx += 1
            exp[s] = nx
    np.testing.assert_array_equal(target, exp)
    assert nxt == nx
    assert (exp[[43, 44]] == dead[0]).all()


# ---------------------------------------------------------------- NIW update
def simple_update_ref(stat, prop, pc, prior_icov, samp, sp_size):
    """calc_simple_update (update_params.cu:204-381) with its float/double
    promotions."""
    n = int(stat[0])
    pc = f32(pc)
    count = float(n)
    total = float(n) + float(f32(pc * f32(50)))
    if prop:  # _add_sigma_smoothing(sample_sigma_shape, count = (int)pc, pc, sp_size)
        ci = int(pc)
        tc = 1.0 * float(f32(f32(ci) + f32(pc * f32(50))))
        pi = np.array([(float(f32(pc * f32(sp_size))) + ci * samp[0]) / (tc + 3.0),
                       (ci * samp[1]) / (tc + 3.0),
                       (float(f32(pc * f32(sp_size))) + ci * samp[2]) / (tc + 3.0)])
    else:
        pi = np.asarray(prior_icov, np.float64)
    det = math.sqrt(pi[0] * pi[2] - pi[1] * pi[1])
    prior_sigma = float(pc) / det * pi
    mu_app = np.array([f32(stat[1 + c] / FX / count) for c in range(3)], np.float32)
    mx, my = stat[4] / count, stat[5] / count
    sx = (stat[6] - count * mx * mx + float(pc) * prior_sigma[0]) / (total + 3.0)
    sy = (stat[7] - count * mx * my + float(pc) * prior_sigma[1]) / (total + 3.0)
    sz = (stat[8] - count * my * my + float(pc) * prior_sigma[2]) / (total + 3.0)
    d = sx * sz - sy * sy
    if d <= 0:
        sx, sz = sx + 1e-5, sz + 1e-5
        d = sx * sz - sy * sy
        if d <= 0:
            d = 1e-5
    return mu_app, np.array([mx, my]), np.array([sz / d, -sy / d, sx / d]), math.log(d)


@pytest.mark.parametrize("prop", [0, 1])
def test_calc_simple_update_both_branches(orc, prop):
    rng = np.random.default_rng(3 + prop)
    for _ in range(200):
        H, W = 20, 24
        m = rng.uniform(0, 1, (H, W)) < rng.uniform(0.05, 0.9)
        if m.sum() < 2:
            continue
        yy, xx = np.mgrid[0:H, 0:W]
        col = rng.uniform(0, 1, (int(m.sum()), 3)).astype(np.float32)
        x, y = xx[m].astype(np.int64), yy[m].astype(np.int64)
        stat = np.array([m.sum()] + [int(np.rint(col[:, c].astype(np.float64) * FX).sum())
                                      for c in range(3)] +
                        [x.sum(), y.sum(), (x * x).sum(), (x * y).sum(), (y * y).sum()],
                        np.int64)
        pc = float(f32(rng.uniform(36, 400)))
        a, c = rng.uniform(0.5, 4, 2)
        icov = np.array([a, rng.uniform(-0.5, 0.5) * math.sqrt(a * c), c])
        samp = np.array([rng.uniform(5, 40), rng.uniform(-3, 3), rng.uniform(5, 40)])
        got = orc.calc_simple_update(stat, prop, pc, icov, samp, 25)
        ref = simple_update_ref(stat, prop, pc, icov, samp, 25)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_allclose(got[1], ref[1], rtol=1e-15)
        np.testing.assert_allclose(got[2], ref[2], rtol=1e-12)
        assert abs(got[3] - ref[3]) <= 1e-12 * max(1, abs(ref[3]))


def test_calc_simple_update_branches_differ(orc):
    stat = np.array([50, 0, 0, 0, 500, 400, 6000, 4100, 3900], np.int64)
    icov = np.array([2.0, 0.0, 2.0])
    samp = np.array([30.0, 4.0, 12.0])
    a = orc.calc_simple_update(stat, 0, 100.0, icov, samp, 25)
    b = orc.calc_simple_update(stat, 1, 100.0, icov, samp, 25)
    assert not np.allclose(a[2], b[2])


# ---------------------------------------------------------------- parameter import
def test_invert_cov_keeps_the_offdiagonal_sign(orc):
    """invert_cov (sparams_io.cu:504-520) divides the adjugate's diagonal by
    the determinant but keeps the off-diagonal's sign: against
    numpy.linalg.inv, entry (0,1) is negated."""
    rng = np.random.default_rng(9)
    for _ in range(100):
        a, c = rng.uniform(0.1, 10, 2)
        b = rng.uniform(-0.9, 0.9) * math.sqrt(a * c)
        inv = np.linalg.inv(np.array([[a, b], [b, c]]))
        got = orc.invert_cov([a, b, c])
        np.testing.assert_allclose(got, [inv[0, 0], -inv[0, 1], inv[1, 1]], rtol=1e-12)


# ---------------------------------------------------------------- flow pooling
def test_flow_pool_means(orc):
    """run_sp_downsample (sp_pooling.cu:27-71): per-label mean flow and pixel
    count; 2^-24 fixed-point sums are within n * 2^-25 of the FP64 sums."""
    rng = np.random.default_rng(11)
    H, W, K = 40, 50, 12
    seg = voronoi_labels(H, W, K, 3)
    seg[0, :5] = -1  # holes are skipped
    flow = rng.normal(0, 3, (H, W, 2)).astype(np.float32)
    ds, cnt = orc.flow_pool(flow, seg, K + 2)
    for k in range(K + 2):
        m = seg == k
        assert cnt[k] == m.sum()
        for c in range(2):
            if cnt[k] == 0:
                assert ds[k, c] == 0
                continue
            mean = flow[m][:, c].astype(np.float64).mean()
            assert abs(float(ds[k, c]) - mean) <= 2.0 ** -25 + abs(mean) * 2.0 ** -23


# ---------------------------------------------------------------- whole split path
def test_split_eval_matches_bass_split(orc):
    """The unit split on the labels of a BASS result proposes the same halves
    as the independent BFS for every label (both directions)."""
    img = synth.voronoi_image(48, 60, 7, seed=12)
    seg, p = orc.bass(img, orc.make_cfg(niters=8, sp_size=9, potts=10.0))
    K = int(seg.max()) + 1
    for direction in (1, 2):
        halves, _, _, _ = orc.split_eval(img, seg, K, direction, 8.1e-5)
        h_r, _, _ = split_ref(img, seg, K, direction, 8.1e-5)
        np.testing.assert_array_equal(halves, h_r)


# ------------------------------------------------- shape (inverse-Wishart) term
def shape_ref(pts, psi, nu):
    """marginal_likelihood_shape_p (split_merge_prop.cu:1596-1621) from the
    pixel coordinates themselves: numpy's biased covariance and det."""
    pts = np.asarray(pts, np.float64)
    cov = np.cov(pts.T, bias=True)
    det_est = abs(np.linalg.det(cov))
    if det_est < 0.001:
        return -1e7
    det_pr = abs(psi[0] * psi[2] - psi[1] * psi[1])
    return nu * math.log(det_pr) - (nu + len(pts)) * math.log(det_est)


def sums_of(pts):
    x, y = np.asarray(pts, np.int64).T
    return np.array([x.sum(), y.sum(), (x * x).sum(), (x * y).sum(), (y * y).sum()], np.int64)


def test_shape_marginal_closed_form(orc):
    rng = np.random.default_rng(11)
    for it in range(300):
        n = int(rng.integers(3, 400))
        c = rng.uniform(0, 1900, 2)
        A = rng.normal(0, 1, (2, 2)) * rng.uniform(1, 30)
        pts = np.unique(np.round(c + rng.normal(0, 1, (n, 2)) @ A).astype(np.int64), axis=0)
        psi = np.array([rng.uniform(5, 400), rng.uniform(-20, 20), rng.uniform(5, 400)])
        nu = float(rng.uniform(1, 500))
        got = orc.shape_marginal(len(pts), sums_of(pts), psi, nu)
        exp = shape_ref(pts, psi, nu)
        if exp == -1e7:
            assert got == -1e7
        else:
            assert abs(got - exp) <= 1e-9 * max(1.0, abs(exp)), (it, got, exp)
    # collinear pixels: det 0 -> the reference's -1e7
    line = [(x, 2 * x + 3) for x in range(20)]
    assert orc.shape_marginal(20, sums_of(line), np.array([30.0, 0.0, 30.0]), 10.0) == -1e7


def test_shape_term_changes_the_split(orc):
    """The term is live when on: BASS of small superpixels (sp 8, prior
    scatter det 1/64 above the 0.001 cut) gives other labels with it than
    without it; at sp 40 the prior scatter's det (1/1600) is below the cut,
    the term is 0 and the labels are the same."""
    img = synth.voronoi_image(60, 90, 10, seed=3)
    segs = {}
    for sp in (8, 40):
        for sm in (0, 1):
            cfg = orc.make_cfg(niters=16, sp_size=sp, potts=10.0, shape_marginal=sm)
            segs[sp, sm] = orc.bass(img, cfg)[0]
    assert not np.array_equal(segs[8, 0], segs[8, 1])
    np.testing.assert_array_equal(segs[40, 0], segs[40, 1])
