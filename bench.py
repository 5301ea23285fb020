#!/usr/bin/env python
"""Streaming BIST benchmark: frames/s and p50/p99 frame latency over 16
mixed-resolution streams (4 each at 360x640, 480x854, 720x1280, 1080x1920),
served (batched across streams, continuous batching) versus the plain
one-frame-at-a-time loop.

Inputs are synthetic Voronoi video with per-region affine motion (the
config-1 motion model: 120 regions, |v| <= 4 px/frame, ground-truth flow
plus 0.5 px noise), generated on the GPU and resident in HBM before any
timed region.  Every propagated frame runs flow shift, hole fill,
connectivity and 100 refinement iterations of 4 boundary sweeps with
split/merge every 4 iterations.

Measurements:
  * closed loop (the JSON ``value``): a step is one tick in which every
    stream submits one propagated frame at once and the server drains them;
    ``steps`` ticks are timed after ``warmup`` ticks (frame 0 of every
    stream, BASS, is warm-up).  This is the served capacity.
  * open loop (the JSON ``p50_ms`` / ``p99_ms``): config 3's arrival
    process, each stream a Poisson process of 30 frames/s, frames submitted
    at their arrival times from the main thread while two executors run
    batches; latency is arrival to completion of each propagated frame.
  * the plain loop: the same ticks, one stsp_prop call per frame.
  * the fused sweep kernel against the HBM roofline, and the CPU oracle.

The open loop runs in a child process (this script with --open-loop-child),
which generates the same streams itself; its result is merged into the
JSON line.
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
    ap.add_argument("--sweep-shape", type=int, default=-1,
                    help="fused sweep launch shape (index into core.SWEEP_SHAPES); "
                         "-1: picked by the autotuner (core.SweepTuner)")
    ap.add_argument("--open-frames", type=int, default=60,
                    help="frames per stream of the open-loop (Poisson arrival) run")
    ap.add_argument("--fps", type=float, default=30.0, help="open-loop arrival rate per stream")
    ap.add_argument("--slots", type=int, default=2, help="batch executors of the open loop")
    ap.add_argument("--no-open", action="store_true")
    ap.add_argument("--open-loop-child", action="store_true",
                    help="run only the open loop and print its result (used by the parent)")
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


def ms(x):
    return round(x * 1e3, 3)


def make_streams(args, torch, synth, T):
    streams = []
    for ri, (H, W) in enumerate(RESOLUTIONS):
        for j in range(args.streams_per_res):
            seed = 1000 * ri + j
            vid, flows = synth.voronoi_video_torch(T, H, W, nregions=120, vmax=4.0, seed=seed)
            streams.append({"H": H, "W": W, "vid": vid, "flows": flows})
    torch.cuda.synchronize()
    return streams


def stream_cfg(core, args):
    return core.make_config(niters=args.iters, sp_size=args.sp_size, potts=0.5,
                            prop_niters=args.iters)


def tune_plain(args, streams, torch, core, cfg):
    """Shape of the plain loop's sweeps per resolution: the autotuner's
    probe over one propagated frame of that resolution (a one-frame batch
    in a scratch context; its result is discarded)."""
    tuner = core.SweepTuner()
    runner = core.BatchRunner(1, use_graph=False, tuner=tuner)
    t = max(args.warmup, 1)
    chosen = {}
    for s in streams:
        key = (s["H"], s["W"])
        if key in chosen:
            continue
        ctx = core.Context(s["H"], s["W"], cfg)
        runner.prop([(ctx, s["vid"][t], s["flows"][t - 1], s["spix"], s["params"])])
        torch.cuda.synchronize()
        chosen[key] = ctx.sweep_shape
        core.get_context(s["H"], s["W"], cfg, "cuda").set_tuning(ctx.sweep_shape)
        ctx.close()
    runner.close()
    return {f"{h}x{w}": v for (h, w), v in chosen.items()}


def run_plain(args, streams, torch, core):
    cfg = stream_cfg(core, args)
    core.set_sweep_shape(args.sweep_shape if args.sweep_shape >= 0 else core.DEFAULT_SWEEP_SHAPE)
    for s in streams:
        s["spix"], s["params"] = core.bass_frame(s["vid"][0], cfg)
    for t in range(1, args.warmup + 1):
        for s in streams:
            s["spix"], s["params"], _ = core.prop_frame(s["vid"][t], s["flows"][t - 1], s["spix"],
                                                        s["params"], cfg)
    shapes = tune_plain(args, streams, torch, core, cfg) if args.sweep_shape < 0 else None
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
    return {"elapsed": time.perf_counter() - t0, "lat": lat, "sweep_shapes": shapes}


def run_closed(args, streams, torch, core, timing=True):
    """Closed loop: each tick submits one frame of every stream at once; the
    server drains them (one batch of all streams per tick)."""
    from st_spix_amd import serve
    if args.sweep_shape >= 0:
        core.set_sweep_shape(args.sweep_shape)
    srv = serve.Server(hbm_budget_gb=64, max_batch=len(streams), use_graph=bool(args.graph),
                       slots=1, autotune=args.sweep_shape < 0)
    sids = [srv.open_stream(s["H"], s["W"], stream_cfg(core, args)) for s in streams]
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
    out = {"elapsed": time.perf_counter() - t0, "lat": [r.done - r.arrival for r in srv.completed]}
    log(f"served closed loop: {args.steps} steps done")
    if timing:
        # dominant kernel: the batched fused sweep (all 16 frames per launch)
        # of the last step's batch, launched back to back between two HIP
        # events on the serving stream (no event between launches)
        launches = SWEEP_TIMING_LAUNCHES
        sweep_ms = srv.runner.time_sweeps(launches, stream=srv.cuda_stream)
        search = {}
        for st in srv.streams.values():  # split-proposal searches per path, per frame size
            key = f"{st.H}x{st.W}"
            c = st.ctx.split_search_counts()
            search[key] = {k: search.get(key, {}).get(k, 0) + v for k, v in c.items()}
        px = sum(s["H"] * s["W"] for s in streams)
        shape = next(iter(srv.streams.values())).ctx.sweep_shape  # the last batch's shape
        per_launch = core.SWEEP_SHAPES[shape][2]
        out.update(sweep_ms=sweep_ms, sweep_launches=launches, split_search=search,
                   sweep_bytes=launches * BYTES_PER_PX_SWEEP * per_launch * px,
                   sweep_shape=shape,
                   autotune=srv.tuner.report() if srv.tuner is not None else None)
    srv.close()
    return out


def run_open(args, streams, torch, core):
    """Open loop: every stream a Poisson process of args.fps frames/s; the
    main thread submits each frame at its arrival time while the worker
    threads run batches on args.slots executors."""
    from st_spix_amd import serve
    if args.sweep_shape >= 0:
        core.set_sweep_shape(args.sweep_shape)
    F = min(args.open_frames, streams[0]["vid"].shape[0])
    srv = serve.Server(hbm_budget_gb=64, max_batch=len(streams), use_graph=bool(args.graph),
                       slots=args.slots, autotune=args.sweep_shape < 0)
    sids = [srv.open_stream(s["H"], s["W"], stream_cfg(core, args)) for s in streams]
    arrivals = serve.poisson_arrivals(len(streams), F, args.fps, seed=7)
    srv.start()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for ta, si, f in arrivals:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        s = streams[si]
        srv.submit(sids[si], s["vid"][f], s["flows"][f - 1] if f > 0 else None, arrival=t0 + ta)
    srv.drain()
    t_end = max(r.done for r in srv.completed)
    srv.stop()
    reqs = list(srv.completed)
    frame_of = {}
    for si, sid in enumerate(sids):
        for k, r in enumerate([r for r in reqs if r.sid == sid]):
            frame_of[id(r)] = k
    prop_lat = [r.done - r.arrival for r in reqs if frame_of[id(r)] > 0]
    all_lat = [r.done - r.arrival for r in reqs]
    batches = sum(s.batches for s in srv._slots)
    srv.close()
    return {"frames": len(reqs), "elapsed": t_end - t0, "fps": len(reqs) / (t_end - t0),
            "offered_fps": len(streams) * args.fps, "p50": pct(prop_lat, 50),
            "p99": pct(prop_lat, 99), "p99_all": pct(all_lat, 99), "batches": batches,
            "frames_per_stream": F}


def open_loop_child(args):
    """The open loop in its own process: generates the streams, runs the
    open loop and prints its result as one JSON line."""
    import torch

    from st_spix_amd import core, synth
    streams = make_streams(args, torch, synth, args.open_frames)
    r = run_open(args, streams, torch, core)
    print(json.dumps(r), flush=True)


def run_open_child(args):
    """Runs open_loop_child in a child process; None if it failed."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--open-loop-child",
           "--open-frames", str(args.open_frames), "--fps", str(args.fps), "--slots",
           str(args.slots), "--iters", str(args.iters), "--sp-size", str(args.sp_size),
           "--sweep-shape", str(args.sweep_shape), "--graph", str(args.graph),
           "--streams-per-res", str(args.streams_per_res)]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=sys.stderr, text=True)
    if p.returncode != 0:
        log(f"open loop child failed with status {p.returncode}")
        return None
    return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args):
    """The oracle on the host cores, on a bounded sample of the workload: the
    first propagated frame of every stream (100 iterations, after an
    untimed 20-iteration BASS frame 0), one stream per thread on all the
    cores this process may use; and single-threaded, the first propagated
    frame of one 360x640 and one 480x854 stream."""
    import concurrent.futures as cf

    import oracle
    from st_spix_amd import synth
    threads = max(1, len(os.sched_getaffinity(0)))
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
        return oracle.prop(vid[1], flows[0], seg, {k: v.copy() for k, v in p.items()}, cfg)

    with cf.ThreadPoolExecutor(min(threads, len(jobs))) as ex:
        states = list(ex.map(first, jobs))
        t0 = time.perf_counter()
        list(ex.map(tick, range(len(jobs))))
        dt = time.perf_counter() - t0
    single = [0, args.streams_per_res]  # the first 360x640 and 480x854 streams
    t1 = time.perf_counter()
    for i in single:
        tick(i)
    dt1 = time.perf_counter() - t1
    return {"value": len(jobs) / dt, "unit": "frames/s", "cores": min(threads, len(jobs)),
            "kind": "port",
            "sample": f"oracle (C restatement of the reference) on {cpu_model()} "
                      f"(nproc {os.cpu_count()}, {threads} usable): the first propagated frame "
                      f"of all {len(jobs)} streams ({args.iters} iterations) on "
                      f"{min(threads, len(jobs))} threads in {dt:.2f} s",
            "single_thread": {"frames_per_s": round(len(single) / dt1, 4),
                              "sample": "first propagated frame of one 360x640 and one 480x854 "
                                        f"stream, 1 thread, {dt1:.2f} s"},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "usable_cores": threads}


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
    if args.open_loop_child:
        open_loop_child(args)
        return
    opened = None
    if not args.no_open:  # before this process touches the GPU
        opened = run_open_child(args)
        if opened is not None:
            log(f"open loop: {opened['frames']} frames in {opened['elapsed']:.3f} s, "
                f"p50 {opened['p50'] * 1e3:.1f} ms p99 {opened['p99'] * 1e3:.1f} ms")
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
    streams = make_streams(args, torch, synth, args.warmup + args.steps + 1)
    log(f"{len(streams)} synthetic streams resident in HBM")
    served = run_closed(args, streams, torch, core)
    log(f"served closed loop: {served['elapsed']:.3f} s for {args.steps} steps")
    log(f"split searches per path: {served['split_search']}")
    plain = None if args.no_plain else run_plain(args, streams, torch, core)
    if plain is not None:
        log(f"plain loop: {plain['elapsed']:.3f} s")
    nframes = len(streams) * args.steps
    fps = nframes / served["elapsed"]
    mpx = sum(s["H"] * s["W"] for s in streams) * args.steps / served["elapsed"] / 1e6
    achieved = served["sweep_bytes"] / (served["sweep_ms"] / 1e3) / 1e9 if served["sweep_ms"] else 0
    kernel = "kb_sweep<%d,%d,%d>" % tuple(core.SWEEP_SHAPES[served["sweep_shape"]])
    line = {
        "metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(served["elapsed"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "16 streams (4 each 360x640, 480x854, 720x1280, 1080x1920), "
                               f"propagated frames, {args.iters} iterations x 4 sweeps, "
                               f"sp_size {args.sp_size}, split/merge every 4",
                   "mode": "served closed loop: one batch per tick over all streams"
                           + (", refinement replayed from a hipGraph" if args.graph else ""),
                   "latency": None if opened is None else
                   f"open loop: {args.fps:g} frames/s per stream, Poisson arrivals, "
                   f"{opened['frames_per_stream']} frames per stream, {args.slots} executors",
                   "sweep_shape": served["sweep_shape"],
                   "sweep_shape_from": "autotuner" if args.sweep_shape < 0 else "--sweep-shape"},
        "p50_ms": ms(opened["p50"]) if opened else ms(pct(served["lat"], 50)),
        "p99_ms": ms(opened["p99"]) if opened else ms(pct(served["lat"], 99)),
        "mpix_per_s": round(mpx, 2),
        "closed_loop_latency": {"p50_ms": ms(pct(served["lat"], 50)),
                                "p99_ms": ms(pct(served["lat"], 99))},
        "roofline": {"bound": "hbm",
                     "kernel": kernel + " (fused sweeps over all frames of the step)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": sweep_traffic(kernel),
                     "launches": served["sweep_launches"],
                     "avg_launch_ms": round(served["sweep_ms"] / max(1, served["sweep_launches"]), 4)},
    }
    if opened is not None:
        line["open_loop"] = {"fps": round(opened["fps"], 3), "offered_fps": opened["offered_fps"],
                             "p50_ms": ms(opened["p50"]), "p99_ms": ms(opened["p99"]),
                             "p99_all_frames_ms": ms(opened["p99_all"]),
                             "batches": opened["batches"]}
    if served.get("autotune") is not None:
        line["autotune"] = served["autotune"]
    if plain is not None:
        pfps = nframes / plain["elapsed"]
        line["plain"] = {"frames_per_s": round(pfps, 3),
                         "p50_ms": ms(pct(plain["lat"], 50)), "p99_ms": ms(pct(plain["lat"], 99)),
                         "sweep_shapes": plain["sweep_shapes"]}
        line["served_over_plain"] = round(fps / pfps, 3)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
