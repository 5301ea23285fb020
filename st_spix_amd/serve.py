"""Serving loop: many concurrent video streams, continuous batching.

Each stream keeps its labels and superpixel parameters resident in HBM
between frames (its own stsp_ctx workspace).  Streams are admitted against
an HBM budget.  Requests (one frame of one stream) queue per stream; every
scheduling step takes the oldest pending frame of each stream that has one
(up to ``max_batch``), runs BASS for streams that have no state yet and the
batched propagation + refinement (stsp_prop_batch) for the rest, and records
per-frame latency from arrival to completion.  Frames that arrive while a
batch is on the GPU join the next step, so work from different streams
overlaps on the device instead of queueing behind one another.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional

import torch

from . import _lib, core


class AdmissionError(RuntimeError):
    """Raised when a new stream would exceed the HBM budget."""


def estimate_stream_bytes(H: int, W: int, sp_size: int) -> int:
    """Device bytes one stream holds: pixel buffers, the label workspace
    (sized for BASS, 30 slots per initial superpixel, +25% headroom) and the
    resident labels/parameters."""
    npix = H * W
    K0 = _lib.lib().stsp_bass_capacity(sp_size, H, W) // 30
    cap = int(30 * K0 * 1.25)
    per_label = 900  # parameters, statistics, split/merge state, sweep record
    return npix * 45 + cap * per_label + npix * 4 + 10 * K0 * 200


class Admission:
    """HBM budget bookkeeping of the server (host state only)."""

    def __init__(self, budget_bytes: int):
        self.budget = int(budget_bytes)
        self.used = 0

    def reserve(self, nbytes: int) -> None:
        if self.used + nbytes > self.budget:
            raise AdmissionError(f"stream needs {nbytes} B; "
                                 f"{self.budget - self.used} B of the budget left")
        self.used += nbytes

    def release(self, nbytes: int) -> None:
        self.used = max(0, self.used - nbytes)


@dataclass
class Request:
    sid: int
    img: torch.Tensor
    flow: Optional[torch.Tensor]
    arrival: float
    done: float = math.nan


@dataclass
class Stream:
    sid: int
    H: int
    W: int
    cfg: object
    ctx: core.Context
    nbytes: int
    spix: Optional[torch.Tensor] = None
    params: Optional[dict] = None
    frames: int = 0
    queue: Deque[Request] = field(default_factory=collections.deque)


class Server:
    def __init__(self, hbm_budget_gb: float = 64.0, max_batch: int = 16,
                 use_graph: bool = False, device="cuda"):
        self.device = torch.device(device)
        self.admission = Admission(int(hbm_budget_gb * (1 << 30)))
        self.max_batch = max_batch
        self.streams: Dict[int, Stream] = {}
        self.next_sid = 0
        self.runner = core.BatchRunner(max_batch, use_graph)
        # graph replay needs a non-default stream; default priority
        self.cuda_stream = torch.cuda.Stream(device=self.device)
        self.completed: List[Request] = []

    # -- stream lifecycle -------------------------------------------------
    def open_stream(self, H: int, W: int, cfg=None) -> int:
        cfg = cfg if cfg is not None else core.make_config()
        need = estimate_stream_bytes(H, W, cfg.sp_size)
        self.admission.reserve(need)
        try:
            ctx = core.Context(H, W, cfg, self.device)
        except Exception:
            self.admission.release(need)
            raise
        ctx.set_tuning(core._SWEEP_SHAPE["value"])
        sid = self.next_sid
        self.next_sid += 1
        self.streams[sid] = Stream(sid, H, W, cfg, ctx, need)
        return sid

    def close_stream(self, sid: int) -> None:
        st = self.streams.pop(sid)
        self.admission.release(st.nbytes)
        torch.cuda.synchronize(self.device)
        st.ctx.close()

    # -- requests -----------------------------------------------------------
    def submit(self, sid: int, img: torch.Tensor, flow: Optional[torch.Tensor] = None,
               arrival: Optional[float] = None) -> Request:
        st = self.streams[sid]
        if st.frames == 0 and len(st.queue) == 0:
            flow = None  # the first frame of a stream is a BASS frame
        r = Request(sid, img, flow, time.perf_counter() if arrival is None else arrival)
        st.queue.append(r)
        return r

    def pending(self) -> int:
        return sum(len(s.queue) for s in self.streams.values())

    def step(self) -> List[Request]:
        """Runs one batch: the oldest pending frame of every stream that has
        one (oldest arrivals first, at most max_batch)."""
        heads = sorted((s.queue[0].arrival, s.sid) for s in self.streams.values() if s.queue)
        heads = heads[: self.max_batch]
        if not heads:
            return []
        fresh, warm = [], []
        for _, sid in heads:
            st = self.streams[sid]
            (fresh if st.params is None else warm).append(st)
        done: List[Request] = []
        with torch.cuda.stream(self.cuda_stream):
            if fresh:
                reqs = [st.queue.popleft() for st in fresh]
                outs = self.runner.bass([(st.ctx, r.img) for st, r in zip(fresh, reqs)],
                                        stream=self.cuda_stream)
                for st, r, (spix, params) in zip(fresh, reqs, outs):
                    st.spix, st.params = spix, params
                    st.frames += 1
                    done.append(r)
            if warm:
                items = []
                reqs = []
                for st in warm:
                    r = st.queue.popleft()
                    if r.flow is None:
                        raise RuntimeError(f"stream {st.sid}: frame {st.frames} needs a flow")
                    items.append((st.ctx, r.img, r.flow, st.spix, st.params))
                    reqs.append(r)
                outs = self.runner.prop(items, stream=self.cuda_stream)
                for st, r, (spix, params, _) in zip(warm, reqs, outs):
                    st.spix, st.params = spix, params
                    st.frames += 1
                    done.append(r)
        self.cuda_stream.synchronize()
        t = time.perf_counter()
        for r in done:
            r.done = t
        self.completed.extend(done)
        return done

    def labels(self, sid: int) -> torch.Tensor:
        return self.streams[sid].spix

    def close(self) -> None:
        for sid in list(self.streams):
            self.close_stream(sid)
        self.runner.close()


def latency_percentiles(reqs: List[Request], qs=(50, 99)) -> Dict[int, float]:
    lat = sorted(r.done - r.arrival for r in reqs if not math.isnan(r.done))
    out = {}
    for q in qs:
        if not lat:
            out[q] = math.nan
            continue
        k = min(len(lat) - 1, max(0, int(math.ceil(q / 100.0 * len(lat))) - 1))
        out[q] = lat[k]
    return out
