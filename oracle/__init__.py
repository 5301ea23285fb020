"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker.  The product
library never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

PARAM_FIELDS = [
    # name, dtype, width
    ("mu_app", np.float32, 3),
    ("prior_mu_app", np.float32, 3),
    ("mu_shape", np.float64, 2),
    ("sigma_shape", np.float64, 3),
    ("logdet_sigma_shape", np.float32, 1),
    ("prior_mu_shape", np.float64, 2),
    ("prior_sigma_shape", np.float64, 3),
    ("sample_sigma_shape", np.float64, 3),
    ("counts", np.int32, 1),
    ("sm_counts", np.int32, 1),
    ("prior_counts", np.float32, 1),
    ("ids", np.int32, 1),
    ("invalid_counts", np.float32, 1),
]


class Cfg(ctypes.Structure):
    _fields_ = [("niters", ctypes.c_int), ("niters_seg", ctypes.c_int),
                ("sm_start", ctypes.c_int), ("sp_size", ctypes.c_int),
                ("sm_period", ctypes.c_int), ("prop_niters", ctypes.c_int),
                ("sigma2_app", ctypes.c_float), ("potts", ctypes.c_float),
                ("alpha_hastings", ctypes.c_float), ("split_alpha", ctypes.c_float),
                ("merge_alpha", ctypes.c_float), ("thresh_relabel", ctypes.c_float),
                ("thresh_new", ctypes.c_float), ("target_nspix", ctypes.c_int),
                ("shape_marginal", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n, _, _ in PARAM_FIELDS]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.orc_det_log.restype = ctypes.c_double
        L.orc_det_log.argtypes = [ctypes.c_double]
        L.orc_shape_marginal.restype = ctypes.c_double
        L.orc_shape_marginal.argtypes = [ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_cfg(**kw) -> Cfg:
    d = dict(niters=500, niters_seg=4, sm_start=0, sp_size=25, sm_period=4,
             prop_niters=0, sigma2_app=8.1e-5, potts=10.0,
             alpha_hastings=float(np.log(0.5)), split_alpha=0.0, merge_alpha=0.0,
             thresh_relabel=1e-5, thresh_new=1e-2, target_nspix=0, shape_marginal=0)
    d.update(kw)
    return Cfg(**d)


def alloc_params(n: int) -> dict:
    return {name: np.zeros((n, w) if w > 1 else (n,), dt) for name, dt, w in PARAM_FIELDS}


def params_struct(pd: dict) -> Params:
    return Params(**{name: _p(pd[name]) for name, _, _ in PARAM_FIELDS})


def trim(pd: dict, n: int) -> dict:
    return {k: v[:n].copy() for k, v in pd.items()}


def nspix_from_spsize(sp_size: int, W: int, H: int) -> int:
    return lib().orc_nspix_from_spsize(sp_size, W, H)


def init_hex(H: int, W: int, sp_size: int, brute: bool = False) -> np.ndarray:
    seg = np.zeros((H, W), np.int32)
    fn = lib().orc_init_hex_bruteforce if brute else lib().orc_init_hex
    fn(_p(seg), sp_size, W, H)
    return seg


def simple_lut() -> np.ndarray:
    out = np.zeros(256, np.uint8)
    lib().orc_simple_lut(_p(out))
    return out


def det_log(x: float) -> float:
    return lib().orc_det_log(float(x))


def label_stats(img: np.ndarray, seg: np.ndarray, nlabels: int) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    seg = np.ascontiguousarray(seg, np.int32)
    out = np.zeros((nlabels, 12), np.int64)
    H, W = seg.shape
    lib().orc_label_stats(_p(img), _p(seg), H * W, W, int(nlabels), _p(out))
    return out


def update_seg(img, seg, mu_app, mu_shape, sigma_shape, logdet, counts, niters_seg,
               sigma2_app, potts):
    seg = np.ascontiguousarray(seg, np.int32).copy()
    H, W = seg.shape
    args = [np.ascontiguousarray(a, t) for a, t in
            ((img, np.float32), (mu_app, np.float32), (mu_shape, np.float64),
             (sigma_shape, np.float64), (logdet, np.float64), (counts, np.int32))]
    lib().orc_update_seg(_p(args[0]), _p(seg), H, W, _p(args[1]), _p(args[2]),
                         _p(args[3]), _p(args[4]), _p(args[5]), niters_seg,
                         ctypes.c_float(sigma2_app), ctypes.c_float(potts))
    return seg


def bass(img: np.ndarray, cfg: Cfg):
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape[:2]
    cap = nspix_from_spsize(cfg.sp_size, W, H) * 30
    out = alloc_params(cap)
    seg = np.zeros((H, W), np.int32)
    n = lib().orc_bass(_p(img), H, W, ctypes.byref(cfg), _p(seg),
                       ctypes.byref(params_struct(out)), cap)
    if n < 0:
        raise RuntimeError(f"oracle bass failed: {n}")
    return seg, trim(out, n)


def prop(img, flow, spix_prev, params_prev: dict, cfg: Cfg, niters_state: int = 0,
         want_maps: bool = False):
    """One propagated frame.  params_prev is updated in place (prior_counts,
    invalid_counts), as the reference's shift_and_fill does."""
    img = np.ascontiguousarray(img, np.float32)
    flow = np.ascontiguousarray(flow, np.float32)
    spix_prev = np.ascontiguousarray(spix_prev, np.int32)
    H, W = img.shape[:2]
    nprev = len(params_prev["ids"])
    cap = nprev * 10
    out = alloc_params(cap)
    seg = np.zeros((H, W), np.int32)
    shifted = np.zeros((H, W), np.int32)
    filled = np.zeros((H, W), np.int32)
    ns = ctypes.c_int(niters_state)
    n = lib().orc_prop(_p(img), _p(flow), H, W, _p(spix_prev),
                       ctypes.byref(params_struct(params_prev)), nprev, ctypes.byref(cfg),
                       ctypes.byref(ns), _p(seg), ctypes.byref(params_struct(out)), cap,
                       _p(shifted), _p(filled))
    if n < 0:
        raise RuntimeError(f"oracle prop failed: {n}")
    res = (seg, trim(out, n), ns.value)
    if want_maps:
        res = res + (shifted, filled)
    return res


def shift_labels(spix, flow_ds, sizes):
    spix = np.ascontiguousarray(spix, np.int32)
    H, W = spix.shape
    out = np.zeros((H, W), np.int32)
    lib().orc_shift_labels(_p(spix), _p(np.ascontiguousarray(flow_ds, np.float32)),
                           _p(np.ascontiguousarray(sizes, np.int32)), H, W, _p(out))
    return out


def fill_missing(seg, centers, break_iter=1000):
    seg = np.ascontiguousarray(seg, np.int32).copy()
    H, W = seg.shape
    rc = lib().orc_fill_missing(_p(seg), _p(np.ascontiguousarray(centers, np.float64)),
                                H, W, break_iter)
    return seg, rc


def shape_marginal(n, sums, psi, nu):
    """stsp_shape_marginal: the shape log marginal of one part."""
    sums = np.ascontiguousarray(sums, np.int64)
    psi = np.ascontiguousarray(psi, np.float64)
    return lib().orc_shape_marginal(int(n), sums.ctypes.data, psi.ctypes.data, float(nu))


def prop_shift_labels(spix, flow):
    """prop_cuda.shift_labels for one image: spix (H, W) int32, flow
    (nspix, 2) int32 as (dx, dy).  Returns (shifted, counts)."""
    spix = np.ascontiguousarray(spix, np.int32)
    flow = np.ascontiguousarray(flow, np.int32)
    H, W = spix.shape
    shifted = np.empty((H, W), np.int32)
    counts = np.empty((H, W), np.int32)
    lib().orc_prop_shift_labels(_p(spix), _p(flow), flow.shape[0], H, W, _p(shifted), _p(counts))
    return shifted, counts


def prop_fill_missing(seg, centers, missing, break_iter):
    """prop_cuda.fill_missing for one image: seg (H, W) int32, centers
    (n, 2) float32, missing flat indices.  Returns (filled, rounds,
    negatives left)."""
    seg = np.ascontiguousarray(seg, np.int32).copy()
    centers = np.ascontiguousarray(centers, np.float32)
    missing = np.ascontiguousarray(np.asarray(missing).ravel(), np.int32)
    H, W = seg.shape
    neg = ctypes.c_int()
    rounds = lib().orc_prop_fill_missing(_p(seg), _p(centers), centers.shape[0], _p(missing),
                                         missing.size, H, W, int(break_iter), ctypes.byref(neg))
    return seg, rounds, neg.value


def invalidate_disconnected(seg, nspix):
    seg = np.ascontiguousarray(seg, np.int32).copy()
    H, W = seg.shape
    lib().orc_invalidate_disconnected(_p(seg), H, W, nspix)
    return seg


def bist(vid, flows, cfg: Cfg, video_mode: bool = True):
    """Whole-video oracle in the shape of bist_forward (pyapi.cu:46-195)."""
    T = vid.shape[0]
    out = []
    params = None
    spix = None
    niters_state = cfg.niters
    for t in range(T):
        if t == 0 or not video_mode:
            spix, params = bass(vid[t], cfg)
        else:
            spix, params, niters_state = prop(vid[t], flows[t - 1], spix, params, cfg,
                                              niters_state)
        out.append(spix)
    return np.stack(out)


# ---------------------------------------------------------------- unit steps
def calc_joint(img_px, x, y, mu_app, mu_shape, isig, logdet, count, sigma2_app, neigh_neq,
               potts):
    """calc_joint (update_seg.cu:194-259) of one pixel against one label."""
    L = lib()
    L.orc_calc_joint.restype = ctypes.c_float
    a = [np.ascontiguousarray(v, t) for v, t in ((img_px, np.float32), (mu_app, np.float32),
                                                 (mu_shape, np.float64), (isig, np.float64))]
    return float(L.orc_calc_joint(_p(a[0]), int(x), int(y), _p(a[1]), _p(a[2]), _p(a[3]),
                                  ctypes.c_double(logdet), int(count),
                                  ctypes.c_float(sigma2_app), ctypes.c_float(neigh_neq),
                                  ctypes.c_float(potts)))


def nig_marginal(sigma2_app, n, bn):
    L = lib()
    L.orc_nig_marginal.restype = ctypes.c_double
    b = np.ascontiguousarray(bn, np.float64)
    return float(L.orc_nig_marginal(ctypes.c_float(sigma2_app), ctypes.c_double(n), _p(b)))


def calc_simple_update(stat, prop, prior_count, prior_icov, sample_sigma_shape, sp_size):
    stat = np.ascontiguousarray(stat, np.int64)
    mu_app = np.zeros(3, np.float32)
    mu_shape = np.zeros(2)
    isig = np.zeros(3)
    logdet = ctypes.c_double(0)
    lib().orc_calc_simple_update(_p(stat), int(prop), ctypes.c_float(prior_count),
                                 _p(np.ascontiguousarray(prior_icov, np.float64)),
                                 _p(np.ascontiguousarray(sample_sigma_shape, np.float64)),
                                 int(sp_size), _p(mu_app), _p(mu_shape), _p(isig),
                                 ctypes.byref(logdet))
    return mu_app, mu_shape, isig, logdet.value


def invert_cov(cov):
    out = np.zeros(3)
    lib().orc_invert_cov(_p(np.ascontiguousarray(cov, np.float64)), _p(out))
    return out


def flow_pool(flow, seg, nspix):
    flow = np.ascontiguousarray(flow, np.float32)
    seg = np.ascontiguousarray(seg, np.int32)
    ds = np.zeros((nspix, 2), np.float32)
    cnt = np.zeros(nspix, np.int32)
    lib().orc_flow_pool(_p(flow), _p(seg), seg.size, int(nspix), _p(ds), _p(cnt))
    return ds, cnt


def split_eval(img, seg, K, direction, sigma2_app, split_alpha=0.0, prop=None, shifted=None,
               prop_mode=0):
    """Split proposal and Hastings ratios of labels 0..K-1 (all valid)."""
    img = np.ascontiguousarray(img, np.float32)
    seg = np.ascontiguousarray(seg, np.int32)
    H, W = seg.shape
    pr = np.ascontiguousarray(prop if prop is not None else np.zeros(K), np.int32)
    sh = np.ascontiguousarray(shifted if shifted is not None else seg, np.int32)
    halves = np.zeros((H, W), np.int32)
    hasting = np.zeros(K)
    accept = np.zeros(K, np.int32)
    select = np.zeros(K, np.int32)
    lib().orc_split_eval(_p(img), _p(seg), _p(sh), H, W, int(K), _p(pr), int(direction),
                         ctypes.c_float(sigma2_app), ctypes.c_float(split_alpha),
                         int(prop_mode), _p(halves), _p(hasting), _p(accept), _p(select))
    return halves, hasting, accept, select


def merge_eval(img, seg, K, direction, sigma2_app, alpha_hastings, merge_alpha=0.0, prop=None,
               prop_mode=0):
    img = np.ascontiguousarray(img, np.float32)
    seg = np.ascontiguousarray(seg, np.int32)
    H, W = seg.shape
    pr = np.ascontiguousarray(prop if prop is not None else np.zeros(K), np.int32)
    cand = np.zeros(K, np.int32)
    hasting = np.zeros(K)
    winner = np.zeros(K, np.int32)
    lib().orc_merge_eval(_p(img), _p(seg), H, W, int(K), _p(pr), int(direction),
                         ctypes.c_float(sigma2_app), ctypes.c_float(alpha_hastings),
                         ctypes.c_float(merge_alpha), int(prop_mode), _p(cand), _p(hasting),
                         _p(winner))
    return cand, hasting, winner


def relabel_search(mu_app, living, prior_mu_app, nspix_prev, thresh_relabel, thresh_new,
                   nbuf):
    mu_app = np.ascontiguousarray(mu_app, np.float32)
    living = np.ascontiguousarray(living, np.uint8)
    pm = np.ascontiguousarray(prior_mu_app, np.float32)
    n = len(living)
    target = np.zeros(n, np.int32)
    nxt = lib().orc_relabel_search(_p(mu_app), _p(living), n - 1, _p(pm), int(nspix_prev),
                                   ctypes.c_float(thresh_relabel), ctypes.c_float(thresh_new),
                                   int(nbuf), _p(target))
    return target, nxt
