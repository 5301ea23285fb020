"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Labels and integer fields are compared bit-exactly; the
floating-point parameters are also required to be bit-identical, because
both sides evaluate the same IEEE operation sequence (no FMA contraction,
deterministic log, exact integer statistics)."""
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


def assert_labels(gpu, ref, what):
    g = gpu.cpu().numpy() if torch.is_tensor(gpu) else gpu
    bad = np.argwhere(g != ref)
    assert bad.size == 0, f"{what}: {len(bad)} label mismatches, first at {bad[0]} " \
                          f"gpu={g[tuple(bad[0])]} ref={ref[tuple(bad[0])]}"


def assert_params(gp, rp, what):
    for f in FIELDS:
        g = gp[f].cpu().numpy()
        r = rp[f]
        assert g.shape == r.shape, f"{what}.{f}: shape {g.shape} vs {r.shape}"
        same = (g == r) | (np.isnan(g) & np.isnan(r)) if g.dtype.kind == "f" else (g == r)
        if not same.all():
            i = np.argwhere(~same)[0]
            pytest.fail(f"{what}.{f}: {int((~same).sum())} mismatches, first {i}: "
                        f"gpu={g[tuple(i)]!r} ref={r[tuple(i)]!r}")


def gpu_bass(img, **kw):
    from st_spix_amd import core
    cfg = core.make_config(**kw)
    return core.bass_frame(cuda(img), cfg)


def orc_cfg(orc, **kw):
    m = dict(kw)
    if "alpha" in m:
        m["alpha_hastings"] = m.pop("alpha")
    return orc.make_cfg(**m)


@pytest.mark.parametrize("H,W,sp,niters,nreg,seed", [
    (48, 64, 8, 24, 8, 2), (33, 47, 6, 20, 5, 3), (60, 90, 12, 40, 10, 4),
    (2, 2, 1, 3, 1, 0), (17, 5, 3, 9, 2, 1)])
def test_bass_parity_small(orc, H, W, sp, niters, nreg, seed):
    img = synth.voronoi_image(H, W, nreg, seed=seed)
    kw = dict(niters=niters, sp_size=sp, potts=10.0)
    spix, params = gpu_bass(img, **kw)
    torch.cuda.synchronize()
    seg_r, p_r = orc.bass(img, orc_cfg(orc, **kw))
    assert_labels(spix, seg_r, "labels")
    assert_params(params, p_r, "params")


def test_bass_parity_config0(orc):
    """The reference's single-image case: 321x481, 40 Voronoi regions, sp 25,
    potts 10, 500 iterations with split/merge."""
    img = synth.voronoi_image(321, 481, 40, noise=0.03, seed=0)
    kw = dict(niters=500, sp_size=25, potts=10.0)
    spix, params = gpu_bass(img, **kw)
    torch.cuda.synchronize()
    seg_r, p_r = orc.bass(img, orc_cfg(orc, **kw))
    assert_labels(spix, seg_r, "config0 labels")
    assert_params(params, p_r, "config0 params")


@pytest.mark.parametrize("H,W,sp,niters,nreg,seed", [
    (200, 260, 60, 12, 6, 1),    # boxes over 64 px: two-word / two-chunk wave BFS
    (300, 420, 140, 12, 5, 2),   # boxes over 128 px early on: workgroup band BFS
    (250, 820, 330, 8, 3, 7),    # boxes 129-256 px: band BFS with 2 and 4 words per row
    (200, 1100, 420, 6, 2, 5),   # boxes 257-384 px wide: band BFS with 6 words per row
    (1100, 200, 600, 6, 2, 9),   # boxes over 448 rows: LDS bitset BFS
    (700, 900, 500, 8, 3, 3)])   # two-chunk band BFS, LDS bitsets and global-memory BFS
def test_split_bfs_box_sizes(orc, H, W, sp, niters, nreg, seed):
    """Large superpixels route the split proposal through every BFS path."""
    img = synth.voronoi_image(H, W, nreg, seed=seed)
    kw = dict(niters=niters, sp_size=sp, potts=10.0)
    spix, params = gpu_bass(img, **kw)
    torch.cuda.synchronize()
    seg_r, p_r = orc.bass(img, orc_cfg(orc, **kw))
    assert_labels(spix, seg_r, f"{H}x{W} sp{sp} labels")
    assert_params(params, p_r, f"{H}x{W} sp{sp} params")


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5])
def test_sweep_shapes_bit_identical(orc, shape):
    """Every launch shape of the fused sweep (1, 2 or 4 sweeps per launch,
    different tiles) gives the oracle's labels; odd sizes exercise ragged
    tiles."""
    from st_spix_amd import core
    img = synth.voronoi_image(75, 101, 9, seed=6)
    kw = dict(niters=16, sp_size=9, potts=10.0)
    core.set_sweep_shape(shape)
    try:
        spix, params = gpu_bass(img, **kw)
    finally:
        core.set_sweep_shape(core.DEFAULT_SWEEP_SHAPE)
    seg_r, p_r = orc.bass(img, orc_cfg(orc, **kw))
    assert_labels(spix, seg_r, f"shape{shape}")
    assert_params(params, p_r, f"shape{shape}")


def test_bass_deterministic():
    img = synth.voronoi_image(96, 128, 12, seed=7)
    a, pa = gpu_bass(img, niters=40, sp_size=10, potts=10.0)
    b, pb = gpu_bass(img, niters=40, sp_size=10, potts=10.0)
    assert torch.equal(a, b)
    for f in FIELDS:
        assert torch.equal(pa[f], pb[f]) or torch.allclose(pa[f], pb[f], equal_nan=True)


@pytest.mark.parametrize("prop_niters", [0, 12])
def test_prop_parity_video(orc, prop_niters):
    from st_spix_amd import core
    T, H, W = 5, 64, 80
    vid, flows = synth.voronoi_video(T, H, W, nregions=10, vmax=3.0, seed=11)
    kw = dict(niters=30, sp_size=8, potts=10.0, prop_niters=prop_niters)
    cfg = core.make_config(**kw)
    ocfg = orc_cfg(orc, **kw)
    spix, params = core.bass_frame(cuda(vid[0]), cfg)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    assert_labels(spix, seg_r, "frame0")
    assert_params(params, p_r, "frame0")
    for t in range(1, T):
        spix, params, ns = core.prop_frame(cuda(vid[t]), cuda(flows[t - 1]), spix, params, cfg)
        seg_r, p_r, ns_r = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg)
        torch.cuda.synchronize()
        assert ns == ns_r
        assert_labels(spix, seg_r, f"frame{t}")
        assert_params(params, p_r, f"frame{t}")


def test_bist_forward_matches_oracle(orc):
    import st_spix_amd
    T, H, W = 3, 48, 64
    vid, flows = synth.voronoi_video(T, H, W, nregions=8, vmax=2.5, seed=5)
    out = st_spix_amd.bist_forward(cuda(vid), cuda(flows), 20, 8, 10.0, 8.1e-5,
                                   float(np.log(0.5)), 0.0, True, False)
    ref = orc.bist(vid, flows, orc.make_cfg(niters=20, sp_size=8, potts=10.0))
    assert_labels(out, ref, "bist_forward")
    # image mode: every frame is independent BASS
    out2 = st_spix_amd.bist_forward(cuda(vid), cuda(flows), 20, 8, 10.0, 8.1e-5,
                                    float(np.log(0.5)), 0.0, False, False)
    ref2 = orc.bist(vid, flows, orc.make_cfg(niters=20, sp_size=8, potts=10.0), video_mode=False)
    assert_labels(out2, ref2, "bist_forward image mode")
    # the torch extension module itself, positionally as the reference's
    # bist_cuda.bist_forward, and the keyword-override path (ctypes binding)
    from st_spix_amd import bist_cuda
    out3 = bist_cuda.bist_forward(cuda(vid), cuda(flows), 20, 8, 10.0, 8.1e-5,
                                  float(np.log(0.5)), 0.0, True, False)
    assert_labels(out3, ref, "bist_cuda.bist_forward")
    out4 = st_spix_amd.bist_forward(cuda(vid), cuda(flows), 20, 8, 10.0, 8.1e-5,
                                    float(np.log(0.5)), 0.0, True, False, niters_seg=4)
    assert_labels(out4, ref, "bist_forward with an override")


def test_marked_video():
    import st_spix_amd
    rng = np.random.default_rng(0)
    vid = rng.uniform(-0.2, 1.2, (2, 9, 11, 3)).astype(np.float32)
    spix = rng.integers(0, 3, (2, 9, 11)).astype(np.int32)
    color = [0.0, 0.0, 1.0]
    got = st_spix_amd.get_marked_video(cuda(vid), cuda(spix), color).cpu().numpy()
    exp = np.clip(vid, 0, 1)
    for t in range(2):
        s = spix[t]
        b = np.zeros_like(s, bool)
        b[1:, :] |= s[1:, :] > s[:-1, :]
        b[:-1, :] |= s[:-1, :] > s[1:, :]
        b[:, 1:] |= s[:, 1:] > s[:, :-1]
        b[:, :-1] |= s[:, :-1] > s[:, 1:]
        exp[t][b] = color
    np.testing.assert_array_equal(got, exp)


def test_prop_all_holes_raises():
    """A flow that moves every label out of the frame leaves nothing to fill
    from: the reference exits (fill_missing.cu:97-111); here the call raises."""
    from st_spix_amd import _lib, core
    img = synth.voronoi_image(32, 40, 4, seed=1)
    cfg = core.make_config(niters=8, sp_size=8)
    spix, params = core.bass_frame(cuda(img), cfg)
    flow = np.full((32, 40, 2), 500.0, np.float32)
    with pytest.raises(_lib.StspError) as e:
        core.prop_frame(cuda(img), cuda(flow), spix, params, cfg)
    assert e.value.code == -3


@pytest.mark.parametrize("target", [40, 120])
def test_target_nspix_bass(orc, target):
    """Label-count control (bass.cu:65-105): the alphas are steered every
    second iteration until the count is within 5% of the target."""
    from st_spix_amd import core
    img = synth.voronoi_image(64, 80, 10, seed=3)
    kw = dict(niters=10, sp_size=8, potts=10.0, target_nspix=target)
    spix, params = core.bass_frame(cuda(img), core.make_config(**kw))
    torch.cuda.synchronize()
    seg_r, p_r = orc.bass(img, orc_cfg(orc, **kw))
    assert_labels(spix, seg_r, f"target {target}")
    assert_params(params, p_r, f"target {target}")
    n = len(np.unique(seg_r))
    assert 0.95 * target < n < 1.05 * target


def test_target_nspix_prop(orc):
    """The control in propagated frames (prop.cu:356-401)."""
    from st_spix_amd import core
    vid, flows = synth.voronoi_video(3, 64, 80, nregions=10, vmax=3.0, seed=11)
    kw = dict(niters=10, sp_size=8, potts=10.0, target_nspix=60)
    cfg, ocfg = core.make_config(**kw), orc_cfg(orc, **kw)
    spix, params = core.bass_frame(cuda(vid[0]), cfg)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    for t in range(1, 3):
        spix, params, ns = core.prop_frame(cuda(vid[t]), cuda(flows[t - 1]), spix, params, cfg)
        seg_r, p_r, ns_r = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg)
        torch.cuda.synchronize()
        assert ns == ns_r
        assert_labels(spix, seg_r, f"frame{t}")
        assert_params(params, p_r, f"frame{t}")


def test_batch_rejects_target_mode():
    """The label-count control is a single-frame mode; batches refuse it."""
    from st_spix_amd import _lib, core
    ctx = core.Context(32, 40, core.make_config(sp_size=8, niters=4, target_nspix=10))
    runner = core.BatchRunner(2)
    img = cuda(synth.voronoi_image(32, 40, 4, seed=1))
    with pytest.raises(_lib.StspError) as e:
        runner.bass([(ctx, img)])
    assert e.value.code == -1
    runner.close()
    ctx.close()


def test_config4_full_size_sm_every_2(orc):
    """Config 4's shape: 1080x1920, 400 regions entering and leaving, sp 15,
    split/merge every 2 iterations (sm_period 2), BASS then two propagated
    frames (6 iterations each, so the oracle finishes in seconds)."""
    from st_spix_amd import core
    vid, flows = synth.voronoi_video(3, 1080, 1920, nregions=400, vmax=4.0, churn=0.05, seed=4)
    kw = dict(niters=6, sp_size=15, potts=0.5, sm_period=2, prop_niters=6)
    cfg, ocfg = core.make_config(**kw), orc_cfg(orc, **kw)
    spix, params = core.bass_frame(cuda(vid[0]), cfg)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    assert_labels(spix, seg_r, "frame0")
    assert_params(params, p_r, "frame0")
    for t in range(1, 3):
        spix, params, ns = core.prop_frame(cuda(vid[t]), cuda(flows[t - 1]), spix, params, cfg)
        seg_r, p_r, ns_r = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg)
        torch.cuda.synchronize()
        assert ns == ns_r
        assert_labels(spix, seg_r, f"frame{t}")
        assert_params(params, p_r, f"frame{t}")


@pytest.mark.parametrize("H,W,sp", [(60, 90, 8), (120, 200, 10)])
def test_shape_marginal_extension_matches_oracle(orc, H, W, sp):
    """The default-off shape term (stsp_ctx_set_shape_marginal) on: BASS and
    two propagated frames equal the oracle's with the term on, bit for bit
    (superpixel sizes at which the term is live)."""
    from st_spix_amd import core
    vid, flows = synth.voronoi_video(3, H, W, nregions=10, vmax=3.0, seed=H)
    kw = dict(niters=16, sp_size=sp, potts=10.0, prop_niters=12)
    ocfg = orc_cfg(orc, shape_marginal=1, **kw)
    cfg = core.make_config(**kw)
    ctx = core.Context(H, W, cfg)
    ctx.set_shape_marginal(True)
    spix, params = core.bass_frame(cuda(vid[0]), cfg, ctx=ctx)
    seg_r, p_r = orc.bass(vid[0], ocfg)
    assert_labels(spix, seg_r, "frame 0")
    assert_params(params, p_r, "frame 0")
    for t in range(1, 3):
        spix, params, _ = core.prop_frame(cuda(vid[t]), cuda(flows[t - 1]), spix, params, cfg,
                                          ctx=ctx)
        seg_r, p_r, _ = orc.prop(vid[t], flows[t - 1], seg_r, p_r, ocfg)
        assert_labels(spix, seg_r, f"frame {t}")
        assert_params(params, p_r, f"frame {t}")
    ctx.close()
