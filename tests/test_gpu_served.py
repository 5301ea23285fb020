"""The batched (served) form: each frame's result in a batch equals its
result alone and the oracle's; hipGraph replays are checked on every replay
with eager work between them."""
import numpy as np
import pytest
import torch

from st_spix_amd import core, synth

from test_gpu_parity import assert_labels, assert_params, cuda, orc_cfg

pytestmark = pytest.mark.gpu

STREAMS = [(48, 64, 8, 21), (40, 56, 8, 22), (64, 80, 10, 23), (33, 50, 7, 24)]


def make_videos(T):
    return [synth.voronoi_video(T, H, W, nregions=8, vmax=3.0, seed=seed)
            for H, W, sp, seed in STREAMS]


def run_oracle(orc, vids, kw):
    out = []
    for (H, W, sp, _), (vid, flows) in zip(STREAMS, vids):
        cfg = orc_cfg(orc, sp_size=sp, **kw)
        seg, p = orc.bass(vid[0], cfg)
        frames = [(seg, {k: v.copy() for k, v in p.items()})]
        for t in range(1, vid.shape[0]):
            seg, p, _ = orc.prop(vid[t], flows[t - 1], seg, p, cfg)
            frames.append((seg, {k: v.copy() for k, v in p.items()}))
        out.append(frames)
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_prop_batch_matches_oracle(orc, use_graph):
    T = 4
    kw = dict(niters=12, potts=10.0, prop_niters=10)
    vids = make_videos(T)
    ref = run_oracle(orc, vids, kw)
    ctxs, state = [], []
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for (H, W, sp, _), (vid, flows) in zip(STREAMS, vids):
            cfg = core.make_config(sp_size=sp, **kw)
            ctx = core.Context(H, W, cfg)
            spix, params = core.bass_frame(cuda(vid[0]), cfg, ctx=ctx)
            ctxs.append(ctx)
            state.append((spix, params))
        runner = core.BatchRunner(8, use_graph=use_graph)
        for t in range(1, T):
            items = [(ctxs[i], cuda(vids[i][0][t]), cuda(vids[i][1][t - 1]), state[i][0],
                      state[i][1]) for i in range(len(STREAMS))]
            outs = runner.prop(items, stream=side)
            side.synchronize()
            # unrelated eager work between graph replays
            junk = torch.randn(1 << 16, device="cuda").sum()
            torch.cuda.synchronize()
            del junk
            for i, (spix, params, _) in enumerate(outs):
                assert_labels(spix, ref[i][t][0], f"stream{i} frame{t}")
                assert_params(params, ref[i][t][1], f"stream{i} frame{t}")
                state[i] = (spix, params)
    runner.close()
    for c in ctxs:
        c.close()


def test_server_alone_vs_in_mix():
    """Each stream's labels are the same when it is served alone and when it
    shares batches with other streams."""
    from st_spix_amd import serve
    T = 3
    vids = make_videos(T)
    cfgs = [core.make_config(niters=10, sp_size=sp, potts=10.0, prop_niters=8)
            for (_, _, sp, _) in STREAMS]

    def serve_streams(idx):
        srv = serve.Server(hbm_budget_gb=8, max_batch=8)
        sids = [srv.open_stream(STREAMS[i][0], STREAMS[i][1], cfgs[i]) for i in idx]
        labels = {i: [] for i in idx}
        for t in range(T):
            for i, sid in zip(idx, sids):
                vid, flows = vids[i]
                srv.submit(sid, cuda(vid[t]), cuda(flows[t - 1]) if t > 0 else None)
            while srv.pending():
                srv.step()
            for i, sid in zip(idx, sids):
                labels[i].append(srv.labels(sid).cpu().numpy())
        srv.close()
        return labels

    mix = serve_streams(list(range(len(STREAMS))))
    for i in range(len(STREAMS)):
        alone = serve_streams([i])
        for t in range(T):
            np.testing.assert_array_equal(mix[i][t], alone[i][t])


def test_server_admission_rejects():
    from st_spix_amd import serve
    srv = serve.Server(hbm_budget_gb=0.05, max_batch=4)
    srv.open_stream(64, 64)
    with pytest.raises(serve.AdmissionError):
        for _ in range(100):
            srv.open_stream(1080, 1920)
    srv.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_bass_batch_matches_oracle(orc, use_graph):
    """Config 2's form: images of different sizes refined in shared launches
    (stsp_bass_batch); each equals the oracle's BASS of that image alone.
    Sizes are drawn like config 2's (H in [240, 720], W in [320, 1280]) but
    scaled by 1/8 so the oracle finishes in seconds; two calls in a row
    replay the same batch composition."""
    rng = np.random.default_rng(7)
    sizes = [(int(rng.integers(30, 91)), int(rng.integers(40, 161))) for _ in range(8)]
    runner = core.BatchRunner(8, use_graph=use_graph)
    ctxs = [core.Context(H, W, core.make_config(sp_size=6, niters=16, potts=10.0))
            for H, W in sizes]
    side = torch.cuda.Stream()  # graphs are captured on a non-default stream
    for rep in range(2):
        imgs = [synth.voronoi_image(H, W, 6, seed=100 * rep + i) for i, (H, W) in enumerate(sizes)]
        with torch.cuda.stream(side):
            outs = runner.bass([(c, cuda(im)) for c, im in zip(ctxs, imgs)], stream=side)
        side.synchronize()
        junk = torch.randn(1 << 16, device="cuda").sum()  # other work between replays
        torch.cuda.synchronize()
        del junk
        for (H, W), im, (spix, params) in zip(sizes, imgs, outs):
            seg_r, p_r = orc.bass(im, orc_cfg(orc, sp_size=6, niters=16, potts=10.0))
            assert_labels(spix, seg_r, f"rep{rep} {H}x{W}")
            assert_params(params, p_r, f"rep{rep} {H}x{W}")
    runner.close()
    for c in ctxs:
        c.close()


def test_bass_batch_config2_sizes(orc):
    """Two images at config 2's full sizes (sp 20) in one batch, with graph
    replay requested on the default stream (it runs eagerly there)."""
    sizes = [(240, 320), (720, 1280)]
    runner = core.BatchRunner(4, use_graph=True)
    kw = dict(sp_size=20, niters=12, potts=0.5)
    ctxs = [core.Context(H, W, core.make_config(**kw)) for H, W in sizes]
    imgs = [synth.voronoi_image(H, W, 30, seed=i) for i, (H, W) in enumerate(sizes)]
    outs = runner.bass([(c, cuda(im)) for c, im in zip(ctxs, imgs)])
    torch.cuda.synchronize()
    for (H, W), im, (spix, params) in zip(sizes, imgs, outs):
        seg_r, p_r = orc.bass(im, orc_cfg(orc, **kw))
        assert_labels(spix, seg_r, f"{H}x{W}")
        assert_params(params, p_r, f"{H}x{W}")
    runner.close()
    for c in ctxs:
        c.close()


def test_time_sweeps_leaves_results_and_next_frame_unchanged(orc):
    """stsp_batch_time_sweeps (the benchmark hook) times extra sweep
    launches over the last batch: the results already returned stay as they
    are and the next frame of every stream still equals the oracle's."""
    T = 3
    kw = dict(niters=12, potts=10.0, prop_niters=10)
    vids = make_videos(T)
    ref = run_oracle(orc, vids, kw)
    ctxs, state = [], []
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for (H, W, sp, _), (vid, flows) in zip(STREAMS, vids):
            cfg = core.make_config(sp_size=sp, **kw)
            ctx = core.Context(H, W, cfg)
            spix, params = core.bass_frame(cuda(vid[0]), cfg, ctx=ctx)
            ctxs.append(ctx)
            state.append((spix, params))
        runner = core.BatchRunner(8, use_graph=True)
        for t in range(1, T):
            items = [(ctxs[i], cuda(vids[i][0][t]), cuda(vids[i][1][t - 1]), state[i][0],
                      state[i][1]) for i in range(len(STREAMS))]
            outs = runner.prop(items, stream=side)
            side.synchronize()
            before = [o[0].clone() for o in outs]
            ms = runner.time_sweeps(6, stream=side)
            side.synchronize()
            assert ms > 0.0
            for i, (spix, params, _) in enumerate(outs):
                assert torch.equal(spix, before[i]), f"stream{i} frame{t} changed"
                assert_labels(spix, ref[i][t][0], f"stream{i} frame{t}")
                assert_params(params, ref[i][t][1], f"stream{i} frame{t}")
                state[i] = (spix, params)
    runner.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_prop_batch_mixed_schedules(orc, use_graph):
    """One batch holding frames with different schedules (split/merge every
    2, 3 and 4 iterations; adaptive and fixed iteration counts) is refined
    in groups with separate table slices and graphs; every frame equals the
    oracle's, on three batches in a row with eager work between them."""
    T = 4
    scheds = [dict(sm_period=2, prop_niters=0), dict(sm_period=3, prop_niters=6),
              dict(sm_period=4, prop_niters=0), dict(sm_period=4, prop_niters=9)]
    vids = make_videos(T)
    ctxs, state, refs = [], [], []
    side = torch.cuda.Stream()
    for (H, W, sp, _), (vid, flows), sc in zip(STREAMS, vids, scheds):
        kw = dict(niters=10, sp_size=sp, potts=10.0, **sc)
        cfg = core.make_config(**kw)
        ocfg = orc_cfg(orc, **kw)
        ctx = core.Context(H, W, cfg)
        with torch.cuda.stream(side):
            spix, params = core.bass_frame(cuda(vid[0]), cfg, ctx=ctx)
        seg, p = orc.bass(vid[0], ocfg)
        frames = [(seg, p)]
        for t in range(1, T):
            seg, p, _ = orc.prop(vid[t], flows[t - 1], seg, p, ocfg)
            frames.append((seg, p))
        refs.append(frames)
        ctxs.append(ctx)
        state.append((spix, params))
    runner = core.BatchRunner(8, use_graph=use_graph)
    got = [[None] * T for _ in STREAMS]
    with torch.cuda.stream(side):
        for t in range(1, T):
            items = [(ctxs[i], cuda(vids[i][0][t]), cuda(vids[i][1][t - 1]), state[i][0],
                      state[i][1]) for i in range(len(STREAMS))]
            outs = runner.prop(items, stream=side)
            side.synchronize()
            torch.randn(1 << 16, device="cuda").sum().item()  # eager work between replays
            for i, (spix, params, _) in enumerate(outs):
                assert_labels(spix, refs[i][t][0], f"stream{i} frame{t}")
                state[i] = (spix, params)
                got[i][t] = params
    # prior/invalid counts of a frame are updated in place by the next one
    for i in range(len(STREAMS)):
        for t in range(1, T):
            assert_params(got[i][t], refs[i][t][1], f"stream{i} frame{t}")
    runner.close()
    for c in ctxs:
        c.close()


def test_server_continuous_batching_matches_alone():
    """Worker threads on two executors, frames arriving over time from a
    feeder thread while batches run: every frame's labels equal those of the
    stream served alone, synchronously."""
    import threading
    import time
    from st_spix_amd import serve
    T = 4
    vids = make_videos(T)
    cfgs = [core.make_config(niters=10, sp_size=sp, potts=10.0, prop_niters=0)
            for (_, _, sp, _) in STREAMS]
    dv = [(cuda(v), cuda(f)) for v, f in vids]
    # alone, synchronous
    alone = []
    for i in range(len(STREAMS)):
        srv = serve.Server(hbm_budget_gb=8, max_batch=4, slots=1)
        sid = srv.open_stream(STREAMS[i][0], STREAMS[i][1], cfgs[i])
        frames = []
        for t in range(T):
            srv.submit(sid, dv[i][0][t], dv[i][1][t - 1] if t > 0 else None)
            while srv.pending():
                srv.step()
            frames.append(srv.labels(sid).cpu().numpy())
        srv.close()
        alone.append(frames)
    # mixed, continuous
    srv = serve.Server(hbm_budget_gb=8, max_batch=3, slots=2, use_graph=True)
    sids = [srv.open_stream(H, W, c) for (H, W, _, _), c in zip(STREAMS, cfgs)]
    reqs = {i: [] for i in range(len(STREAMS))}
    srv.start()

    def feed():
        for t in range(T):
            for i, sid in enumerate(sids):
                reqs[i].append(srv.submit(sid, dv[i][0][t], dv[i][1][t - 1] if t > 0 else None))
                time.sleep(0.002 * (i + 1))

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    srv.drain(timeout=120)
    srv.stop()
    for i in range(len(STREAMS)):
        for t in range(T):
            np.testing.assert_array_equal(reqs[i][t].spix.cpu().numpy(), alone[i][t],
                                          err_msg=f"stream {i} frame {t}")
    assert sum(s.batches for s in srv._slots) >= T  # batches ran as frames arrived
    srv.close()


def test_sweep_tuner_picks_shape_and_keeps_results(orc):
    """The sweep-shape autotuner (core.SweepTuner): the first batch of a
    composition runs with the default shape and is probed (six shapes timed,
    one chosen); later batches run with the chosen shape, in a new graph.
    Every frame, before and after the switch, equals the oracle's."""
    T = 4
    kw = dict(niters=12, potts=10.0, prop_niters=10)
    vids = make_videos(T)
    ref = run_oracle(orc, vids, kw)
    tuner = core.SweepTuner(nlaunch=3)
    runner = core.BatchRunner(8, use_graph=True, tuner=tuner)
    side = torch.cuda.Stream()
    ctxs = [core.Context(H, W, core.make_config(sp_size=sp, **kw)) for H, W, sp, _ in STREAMS]
    with torch.cuda.stream(side):
        outs = runner.bass([(c, cuda(v[0][0])) for c, v in zip(ctxs, vids)], stream=side)
        side.synchronize()
        state = list(outs)
        for i, (spix, params) in enumerate(outs):
            assert_labels(spix, ref[i][0][0], f"stream{i} frame0")
            assert_params(params, ref[i][0][1], f"stream{i} frame0")
        for t in range(1, T):
            items = [(ctxs[i], cuda(vids[i][0][t]), cuda(vids[i][1][t - 1]), state[i][0],
                      state[i][1]) for i in range(len(STREAMS))]
            outs = runner.prop(items, stream=side)
            side.synchronize()
            for i, (spix, params, _) in enumerate(outs):
                assert_labels(spix, ref[i][t][0], f"stream{i} frame{t}")
                assert_params(params, ref[i][t][1], f"stream{i} frame{t}")
                state[i] = (spix, params)
    rep = tuner.report()
    assert sorted(r["kind"] for r in rep) == ["bass", "prop"], rep
    for r in rep:
        assert len(r["us_per_sweep"]) == 6 and all(u > 0 for u in r["us_per_sweep"]), r
        assert 0 <= r["shape"] < 6
    prop_shape = [r["shape"] for r in rep if r["kind"] == "prop"][0]
    assert all(c.sweep_shape == prop_shape for c in ctxs)
    runner.close()
    for c in ctxs:
        c.close()
