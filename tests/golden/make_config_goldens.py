"""Writes tests/golden/config_goldens.json: SHA-256 digests of the oracle's
labels and parameters for the configuration workloads of
tests/config_workloads.py, plus digests of their inputs (so a GPU test can
tell whether it rebuilt the same inputs).  The oracle is the C restatement
in oracle/; this script only runs it.  Usage:

    python tests/golden/make_config_goldens.py [--jobs 8] [--only NAME ...]

With --only, the named entries are recomputed and the others kept from the
existing file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import config_workloads as wl  # noqa: E402


def video_goldens(vid, flows, cfg, nframes):
    ocfg = oracle.make_cfg(**cfg)
    seg, p = oracle.bass(vid[0], ocfg)
    frames = [(seg, p)]
    for t in range(1, nframes):
        seg, p, _ = oracle.prop(vid[t], flows[t - 1], seg, p, ocfg)
        frames.append((seg, p))
    # a frame's prior/invalid counts are updated in place by the next frame,
    # so the parameter digests are taken once all frames ran
    return [{"labels": wl.digest(s), "params": wl.params_digest(q), "nspix": int(len(q["ids"]))}
            for s, q in frames]


def config1_job():
    vid, flows = wl.config1_inputs()
    return {"inputs": wl.digest(vid, flows),
            "frames": video_goldens(vid, flows, wl.CONFIG1["cfg"], wl.CONFIG1["T"])}


def config2_job(i):
    img = wl.config2_image(i)
    seg, p = oracle.bass(img, oracle.make_cfg(**wl.CONFIG2["cfg"]))
    return {"shape": list(img.shape[:2]), "inputs": wl.digest(img), "labels": wl.digest(seg),
            "params": wl.params_digest(p), "nspix": int(len(p["ids"]))}


def config3_job(i):
    c = wl.CONFIG3_SAMPLE[i]
    vid, flows = wl.config3_sample_inputs(i, T=c["check"])
    return {"inputs": wl.digest(vid, flows[: c["check"] - 1]),
            "frames": video_goldens(vid, flows, wl.CONFIG3_CFG, c["check"])}


def config3_long_job():
    vid, flows = wl.config3_long_inputs()
    c = wl.CONFIG3_LONG
    return {"inputs": wl.digest(vid, flows),
            "frames": video_goldens(vid, flows, wl.CONFIG3_CFG, c["T"])}


NAMES = ("config1", "config2", "config3_sample", "config3_long")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--only", nargs="*", choices=NAMES, default=None)
    args = ap.parse_args()
    oracle.build()
    path = os.path.join(HERE, "config_goldens.json")
    want = set(args.only or NAMES)
    out = {}
    if args.only and os.path.exists(path):
        out = {k: v for k, v in json.load(open(path)).items() if k not in want}
    t0 = time.time()
    with cf.ProcessPoolExecutor(args.jobs) as ex:
        fut = {}
        if "config1" in want:
            fut["config1"] = ex.submit(config1_job)
        if "config3_long" in want:
            fut["config3_long"] = ex.submit(config3_long_job)
        if "config3_sample" in want:
            fut["config3_sample"] = [ex.submit(config3_job, i)
                                     for i in range(len(wl.CONFIG3_SAMPLE))]
        if "config2" in want:
            fut["config2"] = [ex.submit(config2_job, i) for i in range(wl.CONFIG2["n"])]
        for name, f in fut.items():
            out[name] = [g.result() for g in f] if isinstance(f, list) else f.result()
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
