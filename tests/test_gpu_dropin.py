"""The per-frame drop-ins with the reference's call shapes
(st_spix/prop/__init__.py: indepent_bass 38-85, stream_bass 87-282,
run_prop 284-656; prop_cuda.bass): positional arguments in the reference's
order, its return tuples (pre_fill / post_fill included) and attribute
access on the parameters, all checked against the oracle bit for bit."""
import numpy as np
import pytest
import torch

from st_spix_amd import synth

pytestmark = pytest.mark.gpu

FIELDS = ["mu_app", "prior_mu_app", "mu_shape", "sigma_shape", "logdet_sigma_shape",
          "prior_mu_shape", "prior_sigma_shape", "sample_sigma_shape", "counts",
          "sm_counts", "prior_counts", "ids", "invalid_counts"]


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def check_params(gp, rp, what):
    for f in FIELDS:
        g = getattr(gp, f).cpu().numpy()  # attribute access, as the reference's callers use
        assert np.array_equal(g, rp[f], equal_nan=True), f"{what}.{f}"


def test_run_prop_reference_argument_order(orc):
    from st_spix_amd.bass import bass
    from st_spix_amd.prop import run_prop
    vid, flows = synth.voronoi_video(3, 56, 72, nregions=9, vmax=3.0, seed=13)
    niters, niters_seg, sm_start, sp_size = 10, 4, 0, 8
    sigma2_app, sigma2_size, potts, alpha = 0.01, 0.1, 10.0, -1.3
    spix, params = bass(cuda(vid[0])[None], niters, niters_seg, sm_start, sp_size, sigma2_app,
                        sigma2_size, potts, alpha)
    assert spix.shape == (1, 56, 72)
    ocfg = orc.make_cfg(niters=niters, niters_seg=niters_seg, sm_start=sm_start,
                        sp_size=sp_size, sigma2_app=sigma2_app, potts=potts,
                        alpha_hastings=alpha, prop_niters=niters)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    assert np.array_equal(spix[0].cpu().numpy(), seg_r)
    check_params(params, p_r, "frame0")
    for t in (1, 2):
        # positional, in the order of st_spix/prop/__init__.py:284-285
        out = run_prop(cuda(vid[t])[None], cuda(flows[t - 1])[None], spix, params,
                       niters, niters_seg, sp_size, sigma2_app, sigma2_size, potts, alpha,
                       sm_start)
        assert len(out) == 4
        spix, params, pre_fill, post_fill = out
        seg_r, p_r, _, shifted_r, filled_r = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg,
                                                      want_maps=True)
        torch.cuda.synchronize()
        assert spix.shape == pre_fill.shape == post_fill.shape == (1, 56, 72)
        assert np.array_equal(spix[0].cpu().numpy(), seg_r), f"frame{t} labels"
        assert np.array_equal(pre_fill[0].cpu().numpy(), shifted_r), f"frame{t} pre_fill"
        assert np.array_equal(post_fill[0].cpu().numpy(), filled_r), f"frame{t} post_fill"
        assert (post_fill >= 0).all()
        check_params(params, p_r, f"frame{t}")


def test_stream_bass_returns_reference_tuple(orc):
    from st_spix_amd.prop import stream_bass
    T, H, W = 3, 48, 64
    vid, flows = synth.voronoi_video(T, H, W, nregions=8, vmax=2.5, seed=17)
    sp_size = 8
    spix, plist, pre_fill, post_fill = stream_bass(cuda(vid), cuda(flows), sp_size=sp_size)
    torch.cuda.synchronize()
    assert spix.shape == (T, H, W) and len(plist) == T
    assert pre_fill.shape == post_fill.shape == (T - 1, H, W)
    # the reference's loop: frame 0 with niters = sp_size and its defaults
    # (sigma2_app 0.01, alpha 0.1, potts 10); later frames with alpha -1.3
    kw = dict(niters=sp_size, niters_seg=4, sm_start=0, sp_size=sp_size, sigma2_app=0.01,
              potts=10.0)
    seg_r, p_r = orc.bass(vid[0], orc.make_cfg(alpha_hastings=0.1, **kw))
    assert np.array_equal(spix[0].cpu().numpy(), seg_r)
    pcfg = orc.make_cfg(alpha_hastings=-1.3, prop_niters=sp_size, **kw)
    plist_r = [p_r]
    for t in range(1, T):
        seg_r, p_r, _, sh, fi = orc.prop(vid[t], flows[t - 1], seg_r, p_r, pcfg, want_maps=True)
        plist_r.append(p_r)
        assert np.array_equal(spix[t].cpu().numpy(), seg_r), f"frame{t}"
        assert np.array_equal(pre_fill[t - 1].cpu().numpy(), sh)
        assert np.array_equal(post_fill[t - 1].cpu().numpy(), fi)
    # each frame's prior and invalid counts are updated in place by the next
    # frame's propagation (shift_and_fill), so compare once every frame ran
    for t in range(T):
        check_params(plist[t], plist_r[t], f"frame{t}")


def test_indepent_bass_overrides(orc):
    from st_spix_amd.prop import indepent_bass
    vid, _ = synth.voronoi_video(2, 40, 56, nregions=6, vmax=2.0, seed=19)
    sp_size = 7
    out = indepent_bass(cuda(vid), sp_size=sp_size)
    torch.cuda.synchronize()
    cfg = orc.make_cfg(niters=sp_size, niters_seg=4, sm_start=0, sp_size=sp_size,
                       sigma2_app=0.1, potts=10.0, alpha_hastings=0.01)
    for t in range(2):
        seg_r, _ = orc.bass(vid[t], cfg)
        assert np.array_equal(out[t].cpu().numpy(), seg_r)
