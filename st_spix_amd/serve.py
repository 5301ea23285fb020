"""Serving loop: many concurrent video streams, continuous batching.

Each stream keeps its labels and superpixel parameters resident in HBM
between frames (its own stsp_ctx workspace).  Streams are admitted against
an HBM budget, and the device memory a stream's context actually grows to is
charged to that budget after every frame.  Requests (one frame of one
stream) queue per stream.  A stream has at most one frame on the GPU, since
its next frame starts from that frame's labels.

Scheduling is continuous: ``slots`` batch executors (each a BatchRunner on
its own HIP stream, all of default priority) are driven by worker threads.
Whenever an executor is free it takes the oldest pending frame of every
stream that has one and is not already on the GPU (at most ``max_batch``),
runs BASS for streams without state and the batched propagation +
refinement (stsp_prop_batch) for the rest, and records each frame's latency
from its arrival to its completion.  With ``autotune`` the executors share a
sweep-shape tuner (core.SweepTuner), which probes the launch shapes once per
batch composition.  Frames that arrive while batches run
are admitted into the next free executor's batch, so the work of
consecutive batches overlaps on the device.  The library call releases the
Python interpreter lock, so executors wait on the GPU concurrently.

``step()`` runs one batch synchronously in the caller's thread (for callers
that drive the loop themselves); ``start()`` / ``drain()`` / ``stop()`` run
the worker threads.  The results of a frame are the same whichever batch
and executor it runs in (tests/test_gpu_served.py).
"""
from __future__ import annotations

import collections
import ctypes
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Deque, Dict, List, Optional

import torch

from . import _lib, core


class AdmissionError(RuntimeError):
    """Raised when a new stream would exceed the HBM budget."""


def estimate_stream_bytes(H: int, W: int, sp_size: int) -> int:
    """Device bytes reserved for one stream at admission: pixel buffers, the
    label workspace (sized for BASS, 30 slots per initial superpixel, +25%
    headroom, which also covers a propagated frame's 10 slots per previous
    superpixel up to 3.75 superpixels per initial one) and the resident
    labels/parameters.  Growth beyond it is charged as it happens."""
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

    def grow(self, nbytes: int) -> None:
        """Charges memory a running stream already took (never refused)."""
        self.used += nbytes

    def release(self, nbytes: int) -> None:
        self.used = max(0, self.used - nbytes)


@dataclass
class Request:
    sid: int
    img: torch.Tensor
    flow: Optional[torch.Tensor]
    arrival: float
    start: float = math.nan
    done: float = math.nan
    spix: Optional[torch.Tensor] = None   # the frame's labels once done
    params: Optional[dict] = None


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
    busy: bool = False      # a frame of this stream is on the GPU
    closing: bool = False


class _Slot:
    def __init__(self, max_batch: int, use_graph: bool, device, tuner):
        self.runner = core.BatchRunner(max_batch, use_graph, tuner=tuner, defer_probe=True)
        self.stream = torch.cuda.Stream(device=device)  # default priority
        self.batches = 0


class Server:
    def __init__(self, hbm_budget_gb: float = 64.0, max_batch: int = 16,
                 use_graph: bool = False, device="cuda", slots: int = 2,
                 autotune: bool = False):
        self.device = torch.device(device)
        self.admission = Admission(int(hbm_budget_gb * (1 << 30)))
        self.max_batch = max_batch
        self.streams: Dict[int, Stream] = {}
        self.next_sid = 0
        self.completed: List[Request] = []
        self._cv = threading.Condition()
        # one sweep-shape tuner shared by the executors (core.SweepTuner)
        self.tuner = core.SweepTuner() if autotune else None
        self._slots = [_Slot(max_batch, use_graph, self.device, self.tuner)
                       for _ in range(max(1, slots))]
        self._workers: List[threading.Thread] = []
        self._running = False
        self._error: Optional[BaseException] = None

    # compatibility with callers of the single-executor form
    @property
    def runner(self) -> core.BatchRunner:
        return self._slots[0].runner

    @property
    def cuda_stream(self) -> torch.cuda.Stream:
        return self._slots[0].stream

    # -- stream lifecycle -------------------------------------------------
    def open_stream(self, H: int, W: int, cfg=None) -> int:
        cfg = cfg if cfg is not None else core.make_config()
        need = estimate_stream_bytes(H, W, cfg.sp_size)
        with self._cv:
            self.admission.reserve(need)
        try:
            ctx = core.Context(H, W, cfg, self.device)
            ctx.set_tuning(core._SWEEP_SHAPE["value"])
        except Exception:
            with self._cv:
                self.admission.release(need)
            raise
        with self._cv:
            sid = self.next_sid
            self.next_sid += 1
            self.streams[sid] = Stream(sid, H, W, cfg, ctx, need)
        return sid

    def close_stream(self, sid: int) -> None:
        """Ends a stream once its queued frames are done; frees its budget."""
        with self._cv:
            st = self.streams[sid]
            st.closing = True
            while st.busy or (st.queue and self._running):
                self._cv.wait()
            self.streams.pop(sid)
            self.admission.release(st.nbytes)
        torch.cuda.synchronize(self.device)
        st.ctx.close()

    # -- requests -----------------------------------------------------------
    def submit(self, sid: int, img: torch.Tensor, flow: Optional[torch.Tensor] = None,
               arrival: Optional[float] = None) -> Request:
        with self._cv:
            st = self.streams[sid]
            if st.closing:
                raise RuntimeError(f"stream {sid} is closing")
            if st.frames == 0 and len(st.queue) == 0 and not st.busy:
                flow = None  # the first frame of a stream is a BASS frame
            elif flow is None:
                raise RuntimeError(f"stream {sid}: frame {st.frames + len(st.queue)} needs a flow")
            r = Request(sid, img, flow, time.perf_counter() if arrival is None else arrival)
            st.queue.append(r)
            self._cv.notify_all()
        return r

    def pending(self) -> int:
        """Frames queued or on the GPU."""
        with self._cv:
            return sum(len(s.queue) + int(s.busy) for s in self.streams.values())

    # -- scheduling ---------------------------------------------------------
    def _take_batch(self):
        """Under the lock: the oldest pending frame of every idle stream."""
        heads = sorted((s.queue[0].arrival, s.sid) for s in self.streams.values()
                       if s.queue and not s.busy)
        batch = []
        for _, sid in heads[: self.max_batch]:
            st = self.streams[sid]
            st.busy = True
            batch.append((st, st.queue.popleft()))
        return batch

    def _run_batch(self, slot: _Slot, batch) -> None:
        """Outside the lock: BASS for new streams, one batched propagation for
        the rest, on the slot's stream; then the device-memory charge."""
        t = time.perf_counter()
        for _, r in batch:
            r.start = t
        fresh = [(st, r) for st, r in batch if st.params is None]
        warm = [(st, r) for st, r in batch if st.params is not None]
        s = slot.stream
        with torch.cuda.device(self.device), torch.cuda.stream(s):
            if fresh:
                outs = slot.runner.bass([(st.ctx, r.img) for st, r in fresh], stream=s)
                for (st, r), (spix, params) in zip(fresh, outs):
                    st.spix, st.params = spix, params
                    r.spix, r.params = spix, params
            if warm:
                items = [(st.ctx, r.img, r.flow, st.spix, st.params) for st, r in warm]
                outs = slot.runner.prop(items, stream=s)
                for (st, r), (spix, params, _) in zip(warm, outs):
                    st.spix, st.params = spix, params
                    r.spix, r.params = spix, params
        s.synchronize()
        slot.batches += 1
        t = time.perf_counter()
        for _, r in batch:
            r.done = t
        with torch.cuda.device(self.device):
            slot.runner.probe_pending()  # once per batch composition, after the frames are out
        for st, _ in batch:
            held = st.ctx.device_bytes()
            if held > st.nbytes:
                with self._cv:
                    self.admission.grow(held - st.nbytes)
                st.nbytes = held

    def _finish(self, batch) -> None:
        for st, r in batch:
            st.frames += 1
            st.busy = False
            self.completed.append(r)
        self._cv.notify_all()

    def step(self) -> List[Request]:
        """Runs one batch in the caller's thread (executor 0)."""
        with self._cv:
            if self._running:
                raise RuntimeError("step() while the worker threads run")
            batch = self._take_batch()
        if not batch:
            return []
        try:
            self._run_batch(self._slots[0], batch)
        finally:
            with self._cv:
                self._finish(batch)
        return [r for _, r in batch]

    def _worker(self, slot: _Slot) -> None:
        while True:
            with self._cv:
                while self._running and self._error is None:
                    batch = self._take_batch()
                    if batch:
                        break
                    self._cv.wait()
                else:
                    return
            try:
                self._run_batch(slot, batch)
            except BaseException as e:  # surfaced by drain()/stop()
                with self._cv:
                    self._error = e
                    self._finish(batch)
                return
            with self._cv:
                self._finish(batch)

    def start(self) -> None:
        """Starts one worker thread per executor."""
        with self._cv:
            if self._running:
                return
            self._running = True
        self._workers = [threading.Thread(target=self._worker, args=(s,), daemon=True)
                         for s in self._slots]
        for w in self._workers:
            w.start()

    def drain(self, timeout: Optional[float] = None) -> None:
        """Waits until every submitted frame is done."""
        end = None if timeout is None else time.perf_counter() + timeout
        with self._cv:
            while self._error is None and any(s.queue or s.busy for s in self.streams.values()):
                left = None if end is None else end - time.perf_counter()
                if left is not None and left <= 0:
                    raise TimeoutError("frames still pending")
                self._cv.wait(left)
            if self._error is not None:
                raise RuntimeError("a batch failed") from self._error

    def stop(self) -> None:
        with self._cv:
            self._running = False
            self._cv.notify_all()
        for w in self._workers:
            w.join()
        self._workers = []
        if self._error is not None:
            raise RuntimeError("a batch failed") from self._error

    def labels(self, sid: int) -> torch.Tensor:
        return self.streams[sid].spix

    def close(self) -> None:
        if self._running:
            self.drain()
            self.stop()
        for sid in list(self.streams):
            self.close_stream(sid)
        for s in self._slots:
            s.runner.close()


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


def poisson_arrivals(nstreams: int, nframes: int, fps: float, seed: int = 0):
    """Arrival times (seconds from the start) of every frame of every stream:
    a Poisson process of rate ``fps`` per stream (exponential gaps), streams
    starting at independent random phases within one frame interval."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for s in range(nstreams):
        t = rng.uniform(0, 1.0 / fps)
        for f in range(nframes):
            out.append((t, s, f))
            t += rng.exponential(1.0 / fps)
    out.sort()
    return out
