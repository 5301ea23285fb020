#!/usr/bin/env python
"""Streaming BIST benchmark: frames/s and p50/p99 frame latency over 16
mixed-resolution streams (4 each at 360x640, 480x854, 720x1280, 1080x1920),
served (batched across streams) versus the plain one-frame-at-a-time loop.

A step is one tick: every stream advances by one propagated frame (flow
shift + hole fill + connectivity + 100 refinement iterations of 4 boundary
sweeps with split/merge every 4 iterations), on synthetic Voronoi video with
affine region motion (inputs generated on the GPU, resident in HBM before the
timed region).  Frame 0 of every stream (BASS, 100 iterations) is part of the
untimed warm-up.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RESOLUTIONS = [(360, 640), (480, 854), (720, 1280), (1080, 1920)]
METRIC = "frames/s + p99 frame latency, 16 mixed-res streams, served vs plain loop"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PX_SWEEP = 20  # labels r 4 + img r 12 + labels w 4 per sweep
SWEEPS_PER_ITER = 4      # niters_seg
SWEEP_TIMING_LAUNCHES = 50  # back-to-back sweep launches timed for the roofline


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=100, help="refinement iterations per frame")
    ap.add_argument("--streams-per-res", type=int, default=4)
    ap.add_argument("--sp-size", type=int, default=25)
    ap.add_argument("--sweep-shape", type=int, default=4,
                    help="fused sweep launch shape (index into core.SWEEP_SHAPES)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--graph", type=int, default=1, help="replay the refinement as a hipGraph")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run one device copy of this many MiB first (counter calibration)")
    return ap.parse_args()


_T0 = time.perf_counter()


def log(msg):
    """Progress on stderr; stdout carries only the JSON line."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = min(len(xs) - 1, max(0, int(math.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def make_streams(args, torch, synth):
    streams = []
    T = args.warmup + args.steps + 1
    for ri, (H, W) in enumerate(RESOLUTIONS):
        for j in range(args.streams_per_res):
            seed = 1000 * ri + j
            vid, flows = synth.voronoi_video_torch(T, H, W, nregions=120, vmax=4.0, seed=seed)
            streams.append({"H": H, "W": W, "vid": vid, "flows": flows})
    torch.cuda.synchronize()
    return streams


def run_plain(args, streams, torch, core):
    cfg0 = core.make_config(niters=args.iters, sp_size=args.sp_size, potts=0.5)
    cfg = core.make_config(niters=args.iters, sp_size=args.sp_size, potts=0.5,
                           prop_niters=args.iters)
    core.set_sweep_shape(args.sweep_shape)
    for s in streams:
        s["spix"], s["params"] = core.bass_frame(s["vid"][0], cfg0)
    for t in range(1, args.warmup + 1):
        for s in streams:
            s["spix"], s["params"], _ = core.prop_frame(s["vid"][t], s["flows"][t - 1], s["spix"],
                                                        s["params"], cfg)
    for s in streams:
        core.get_context(s["H"], s["W"], cfg, s["vid"].device).enable_timing(True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.warmup + 1, args.warmup + 1 + args.steps):
        for s in streams:
            a = time.perf_counter()
            s["spix"], s["params"], _ = core.prop_frame(s["vid"][t], s["flows"][t - 1], s["spix"],
                                                        s["params"], cfg)
            lat.append(time.perf_counter() - a)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # dominant kernel: the fused boundary sweep, timed with HIP events
    sweep_ms, sweep_bytes, launches = 0.0, 0.0, 0
    seen = set()
    for s in streams:
        key = (s["H"], s["W"])
        if key in seen:
            continue
        seen.add(key)
        ctx = core.get_context(s["H"], s["W"], cfg, s["vid"].device)
        ms, n = ctx.kernel_time(0)
        sweep_ms += ms
        launches += n
        sweep_bytes += n * BYTES_PER_PX_SWEEP * SWEEPS_PER_ITER * s["H"] * s["W"]
        ctx.enable_timing(False)
    return {"elapsed": elapsed, "lat": lat, "sweep_ms": sweep_ms, "sweep_bytes": sweep_bytes,
            "sweep_launches": launches}


def run_served(args, streams, torch, core):
    """Continuous batching: every step takes one pending frame of each stream
    and refines them in shared launches (stsp_prop_batch)."""
    from st_spix_amd import serve
    core.set_sweep_shape(args.sweep_shape)
    srv = serve.Server(hbm_budget_gb=64, max_batch=len(streams), use_graph=bool(args.graph))
    sids = []
    for s in streams:
        cfg = core.make_config(niters=args.iters, sp_size=args.sp_size, potts=0.5,
                               prop_niters=args.iters)
        sids.append(srv.open_stream(s["H"], s["W"], cfg))
    for t in range(0, args.warmup + 1):
        for sid, s in zip(sids, streams):
            srv.submit(sid, s["vid"][t], s["flows"][t - 1] if t > 0 else None)
        while srv.pending():
            srv.step()
    torch.cuda.synchronize()
    srv.completed.clear()
    t0 = time.perf_counter()
    for t in range(args.warmup + 1, args.warmup + 1 + args.steps):
        now = time.perf_counter()
        for sid, s in zip(sids, streams):
            srv.submit(sid, s["vid"][t], s["flows"][t - 1], arrival=now)
        while srv.pending():
            srv.step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lat = [r.done - r.arrival for r in srv.completed]
    # dominant kernel: the batched fused sweep (all 16 frames per launch) of
    # the last step's batch, launched back to back between two HIP events on
    # the serving stream (no event between launches, so no dispatch bubbles)
    launches = SWEEP_TIMING_LAUNCHES
    ms = srv.runner.time_sweeps(launches, stream=srv.cuda_stream)
    search = {}
    for st in srv.streams.values():  # split-proposal searches per path, per frame size
        key = f"{st.H}x{st.W}"
        c = st.ctx.split_search_counts()
        search[key] = {k: search.get(key, {}).get(k, 0) + v for k, v in c.items()}
    px = sum(s["H"] * s["W"] for s in streams)
    per_launch = core.SWEEP_SHAPES[args.sweep_shape][2]
    srv.close()
    return {"elapsed": elapsed, "lat": lat, "sweep_ms": ms,
            "sweep_bytes": launches * BYTES_PER_PX_SWEEP * per_launch * px,
            "sweep_launches": launches, "split_search": search}


def cpu_baseline(args):
    """The oracle on the host cores: one full tick of the stream mix (every
    stream's propagated frame, same iteration count), one stream per thread
    (the oracle has no shared mutable state; ctypes releases the GIL).  Frame
    0 of each stream is an untimed 20-iteration BASS."""
    import concurrent.futures as cf

    import oracle
    from st_spix_amd import synth
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    jobs = []
    for ri, (H, W) in enumerate(RESOLUTIONS):
        for j in range(args.streams_per_res):
            vid, flows = synth.voronoi_video(2, H, W, nregions=120, vmax=4.0, seed=1000 * ri + j)
            jobs.append((vid, flows))
    cfg0 = oracle.make_cfg(niters=20, sp_size=args.sp_size, potts=0.5)
    cfg = oracle.make_cfg(niters=20, sp_size=args.sp_size, potts=0.5, prop_niters=args.iters)

    def first(job):
        return oracle.bass(job[0][0], cfg0)

    def tick(i):
        vid, flows = jobs[i]
        seg, p = states[i]
        return oracle.prop(vid[1], flows[0], seg, p, cfg)

    with cf.ThreadPoolExecutor(threads) as ex:
        states = list(ex.map(first, jobs))
        t0 = time.perf_counter()
        list(ex.map(tick, range(len(jobs))))
        dt = time.perf_counter() - t0
    return {"value": len(jobs) / dt, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C restatement of the reference), {threads} threads, one stream "
                      f"per thread: one tick of all {len(jobs)} streams (a propagated frame each, "
                      f"{args.iters} iterations) in {dt:.2f} s"}


def sweep_traffic(kernel):
    """HBM bytes per launch of the fused sweep over the same 16 frames, as
    measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by a copy kernel)
    and recorded in profiles/kb_sweep_traffic.md; None if absent or if it
    was measured for another launch shape."""
    path = os.path.join(ROOT, "profiles", "kb_sweep_traffic.md")
    fields = {}
    try:
        for line in open(path):
            if ":" in line and not line.startswith(("#", "|", " ", "`")):
                key, val = line.split(":", 1)
                fields[key.strip()] = val.strip()
    except OSError:
        return None
    if fields.get("kernel") != kernel or "traffic_bytes_per_launch" not in fields:
        return None
    return float(fields["traffic_bytes_per_launch"])


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    import torch

    from st_spix_amd import core, synth
    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written")
        del src, dst
    streams = make_streams(args, torch, synth)
    log(f"{len(streams)} synthetic streams resident in HBM")
    served = run_served(args, streams, torch, core)
    log(f"served: {served['elapsed']:.3f} s for {args.steps} steps")
    log(f"split searches per path: {served['split_search']}")
    plain = None if args.no_plain else run_plain(args, streams, torch, core)
    if plain is not None:
        log(f"plain loop: {plain['elapsed']:.3f} s")
    nframes = len(streams) * args.steps
    fps = nframes / served["elapsed"]
    mpx = sum(s["H"] * s["W"] for s in streams) * args.steps / served["elapsed"] / 1e6
    achieved = served["sweep_bytes"] / (served["sweep_ms"] / 1e3) / 1e9 if served["sweep_ms"] else 0
    kernel = "kb_sweep<%d,%d,%d>" % tuple(core.SWEEP_SHAPES[args.sweep_shape])
    line = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(served["elapsed"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "16 streams (4 each 360x640, 480x854, 720x1280, 1080x1920), "
                               f"propagated frames, {args.iters} iterations x 4 sweeps, "
                               f"sp_size {args.sp_size}, split/merge every 4",
                   "mode": "served: one batch per step over all streams"
                           + (", refinement replayed from a hipGraph" if args.graph else ""),
                   "sweep_shape": args.sweep_shape},
        "p50_ms": round(pct(served["lat"], 50) * 1e3, 3),
        "p99_ms": round(pct(served["lat"], 99) * 1e3, 3),
        "mpix_per_s": round(mpx, 2),
        "roofline": {"bound": "hbm",
                     "kernel": kernel + " (fused sweeps over all frames of the step)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": sweep_traffic(kernel),
                     "launches": served["sweep_launches"],
                     "avg_launch_ms": round(served["sweep_ms"] / max(1, served["sweep_launches"]), 4)},
    }
    if plain is not None:
        pfps = nframes / plain["elapsed"]
        line["plain"] = {"frames_per_s": round(pfps, 3),
                         "p50_ms": round(pct(plain["lat"], 50) * 1e3, 3),
                         "p99_ms": round(pct(plain["lat"], 99) * 1e3, 3)}
        line["served_over_plain"] = round(fps / pfps, 3)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
