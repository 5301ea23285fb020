"""GPU parity of the split proposal's search paths.

The proposal (calc_split_candidate / calc_seg_split, split_merge.cu:757-860;
split_merge_prop.cu:277-323, 1848-1940) is searched by one wavefront for
small bounding boxes, by a workgroup over the whole box, by a workgroup in a
window around the first seed (exact when the seed's connected component ends
inside the window), or over the whole box after the component leaves the
window.  Every path must give the oracle's labels bit for bit; the search
limits are shrunk here so that frames of a few hundred pixels reach every
path, and the per-path counters show that they did.
"""
import numpy as np
import pytest
import torch

from st_spix_amd import synth

pytestmark = pytest.mark.gpu

FIELDS = ["mu_app", "mu_shape", "sigma_shape", "logdet_sigma_shape", "counts",
          "prior_counts", "ids", "invalid_counts", "sample_sigma_shape"]


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def same_params(gp, rp):
    for f in FIELDS:
        g = gp[f].cpu().numpy()
        if g.shape != rp[f].shape or not np.array_equal(g, rp[f], equal_nan=True):
            return f
    return None


def run_video(orc, T, H, W, nreg, sp, niters, seed, search, churn=0.0):
    from st_spix_amd import core
    vid, flows = synth.voronoi_video(T, H, W, nregions=nreg, vmax=3.0, churn=churn, seed=seed)
    kw = dict(niters=niters, sp_size=sp, potts=10.0, prop_niters=niters)
    cfg = core.make_config(**kw)
    ocfg = orc.make_cfg(**kw)
    ctx = core.Context(H, W, cfg)
    ctx.set_split_search(*search)
    spix, params = core.bass_frame(cuda(vid[0]), cfg, ctx=ctx)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    torch.cuda.synchronize()
    assert np.array_equal(spix.cpu().numpy(), seg_r), "frame 0 labels"
    assert same_params(params, p_r) is None, "frame 0 params"
    for t in range(1, T):
        spix, params, ns = core.prop_frame(cuda(vid[t]), cuda(flows[t - 1]), spix, params, cfg,
                                           ctx=ctx)
        seg_r, p_r, ns_r = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg)
        torch.cuda.synchronize()
        bad = np.argwhere(spix.cpu().numpy() != seg_r)
        assert bad.size == 0, f"frame {t}: {len(bad)} label mismatches, first at {bad[0]}"
        f = same_params(params, p_r)
        assert f is None, f"frame {t}: params.{f} differ"
    counts = ctx.split_search_counts()
    ctx.close()
    return counts


@pytest.mark.parametrize("search", [(1, 16, 64), (1, 8, 32), (16, 24, 64), (128, 448, 384),
                                    (16, 448, 384)])
def test_split_search_paths_prop(orc, search):
    """Propagated video (relabelling can leave a label in several pieces, so
    windows that hold the seed's component occur beside windows it leaves)."""
    counts = run_video(orc, 6, 96, 128, 14, 10, 24, seed=21, search=search)
    print("search", search, counts)
    if search[0] == 1:
        assert counts["wave"] == 0
        assert counts["window"] + counts["whole"] > 0
    else:
        assert counts["wave"] > 0


def test_split_search_paths_churn(orc):
    """Regions entering and leaving (config 4's motion) with small windows."""
    counts = run_video(orc, 5, 90, 150, 30, 8, 16, seed=5, search=(1, 12, 64), churn=0.1)
    print("churn", counts)
    assert counts["window"] + counts["whole"] > 0


@pytest.mark.parametrize("search", [(16, 448, 384), (16, 896, 256)])
def test_split_search_large_boxes(orc, search):
    """Few, large superpixels: boxes of a few hundred pixels, searched whole by
    the 8-wave kernel in both orientations (1 to 6 words per band row).  With
    the narrow window (256 columns) the boxes wider than it are searched
    whole and transposed by the big-box kernel, in one or two bands per wave
    (box widths up to 448, then up to 480)."""
    counts = run_video(orc, 4, 256, 480, 6, 90, 20, seed=8, search=search)
    print("large", search, counts)
    assert counts["band"] > 0
