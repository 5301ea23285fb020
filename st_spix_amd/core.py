"""Torch-facing host code over the C ABI: contexts, parameter sets and the
per-frame calls.  PyTorch supplies device memory and the current stream; all
compute runs in libstspix.so.
"""
from __future__ import annotations

import ctypes
import math
import threading
from typing import Dict, Optional, Tuple

import torch

from . import _lib

class ParamSet(dict):
    """Superpixel parameters: a dict of CUDA tensors named after the
    reference's SuperpixelParams fields (st_spix/csrc/bist/structs.h:13-48),
    also readable and writable as attributes (``params.mu_app``), as the
    reference's callers use them (st_spix/prop/__init__.py:175, 300)."""

    def __getattr__(self, name):
        try:
            return self[name]
        except KeyError:
            raise AttributeError(name) from None

    def __setattr__(self, name, value):
        self[name] = value


Params = Dict[str, torch.Tensor]

_DTYPES = {"float32": torch.float32, "float64": torch.float64, "int32": torch.int32}


def _check_input(t: torch.Tensor, name: str) -> None:
    # same contract as the reference's CHECK_INPUT (pyapi.cu:5-7)
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def make_config(niters: int = 500, sp_size: int = 25, potts: float = 0.5,
                sigma2_app: float = 8.1e-5, alpha: float = math.log(0.5),
                split_alpha: float = 0.0, merge_alpha: float = 0.0, niters_seg: int = 4,
                sm_start: int = 0, sm_period: int = 4, prop_niters: int = 0,
                thresh_relabel: float = 1e-5, thresh_new: float = 1e-2,
                target_nspix: int = 0) -> _lib.Config:
    return _lib.Config(niters=niters, niters_seg=niters_seg, sm_start=sm_start,
                       sp_size=sp_size, sm_period=sm_period, prop_niters=prop_niters,
                       sigma2_app=sigma2_app, potts=potts, alpha_hastings=alpha,
                       split_alpha=split_alpha, merge_alpha=merge_alpha,
                       thresh_relabel=thresh_relabel, thresh_new=thresh_new,
                       target_nspix=target_nspix)


class Context:
    """Owns the device workspace for one frame size (stsp_ctx)."""

    def __init__(self, height: int, width: int, cfg: _lib.Config, device=None):
        self.H, self.W = int(height), int(width)
        self.device = torch.device(device if device is not None else "cuda")
        self.cfg = cfg
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stsp_ctx_create(ctypes.byref(cfg), self.H, self.W,
                                                  ctypes.byref(h)), "stsp_ctx_create")
        self.handle = h
        self.sweep_shape = DEFAULT_SWEEP_SHAPE

    def set_tuning(self, sweep_shape: int) -> None:
        _lib.check(_lib.lib().stsp_ctx_set_tuning(self.handle, int(sweep_shape)),
                   "stsp_ctx_set_tuning")
        self.sweep_shape = int(sweep_shape)

    def set_shape_marginal(self, on: bool) -> None:
        """Adds the shape (inverse-Wishart) term to the split ratio
        (stsp_ctx_set_shape_marginal; an extension, off by default)."""
        _lib.check(_lib.lib().stsp_ctx_set_shape_marginal(self.handle, int(bool(on))),
                   "stsp_ctx_set_shape_marginal")

    def set_split_search(self, wave_box: int = 128, win_rows: int = 896,
                         win_cols: int = 512) -> None:
        """Limits of the split proposal's search paths (stsp_ctx_set_split_search);
        results are the same for every setting."""
        _lib.check(_lib.lib().stsp_ctx_set_split_search(self.handle, int(wave_box),
                                                        int(win_rows), int(win_cols)),
                   "stsp_ctx_set_split_search")

    def split_search_counts(self) -> Dict[str, int]:
        """Cumulative split-proposal searches per path (waits for the device)."""
        out = (ctypes.c_longlong * 4)()
        _lib.check(_lib.lib().stsp_ctx_split_search_counts(self.handle, out),
                   "stsp_ctx_split_search_counts")
        return dict(zip(("wave", "band", "window", "whole"), (int(v) for v in out)))

    def device_bytes(self) -> int:
        """Device memory the context holds (stsp_ctx_device_bytes)."""
        n = ctypes.c_size_t(0)
        _lib.check(_lib.lib().stsp_ctx_device_bytes(self.handle, ctypes.byref(n)),
                   "stsp_ctx_device_bytes")
        return int(n.value)

    def enable_timing(self, on: bool = True) -> None:
        _lib.check(_lib.lib().stsp_ctx_enable_timing(self.handle, int(on)),
                   "stsp_ctx_enable_timing")

    def kernel_time(self, kclass: int):
        ms = ctypes.c_double(0)
        n = ctypes.c_longlong(0)
        _lib.check(_lib.lib().stsp_ctx_kernel_time(self.handle, kclass, ctypes.byref(ms),
                                                   ctypes.byref(n)), "stsp_ctx_kernel_time")
        return ms.value, n.value

    def set_config(self, cfg: _lib.Config) -> None:
        self.cfg = cfg
        _lib.check(_lib.lib().stsp_ctx_set_config(self.handle, ctypes.byref(cfg)),
                   "stsp_ctx_set_config")

    def close(self) -> None:
        if self.handle:
            _lib.lib().stsp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CTX: Dict[tuple, Context] = {}

# launch shapes of the fused sweep (tile width, height, sweeps per launch),
# index = stsp_ctx_set_tuning's sweep_shape; the library default is 4
SWEEP_SHAPES = [(32, 16, 1), (32, 32, 2), (64, 32, 4), (64, 64, 2), (64, 32, 2), (128, 64, 2)]
DEFAULT_SWEEP_SHAPE = 4
_SWEEP_SHAPE = {"value": DEFAULT_SWEEP_SHAPE}


def set_sweep_shape(shape: int) -> None:
    """Launch shape of the fused boundary sweep for all contexts (0..5; see
    stsp_ctx_set_tuning in include/stspix.h)."""
    _SWEEP_SHAPE["value"] = int(shape)
    for ctx in _CTX.values():
        ctx.set_tuning(int(shape))


def get_context(height: int, width: int, cfg: _lib.Config, device) -> Context:
    dev = torch.device(device)
    key = (int(height), int(width), dev.index)
    ctx = _CTX.get(key)
    if ctx is None:
        ctx = Context(height, width, cfg, dev)
        ctx.set_tuning(_SWEEP_SHAPE["value"])
        _CTX[key] = ctx
    else:
        ctx.set_config(cfg)
    return ctx


def alloc_params(cap: int, device, zero: bool = True) -> Params:
    """Parameter arrays of `cap` slots.  The export writes every field of
    slots [0, nspix) and results are trimmed to nspix, so callers on the
    served path skip the zero fill (zero=False)."""
    out = ParamSet()
    make = torch.zeros if zero else torch.empty
    for name, dt, w in _lib.PARAM_FIELDS:
        shape = (cap, w) if w > 1 else (cap,)
        out[name] = make(shape, dtype=_DTYPES[dt], device=device)
    return out


_ITEMSIZE = {"float32": 4, "float64": 8, "int32": 4}
_ALIGN = 256


class _OutBatch:
    """The outputs of a batched call (each frame's label map and parameter
    arrays of `cap` slots) carved from one device allocation, 256-byte
    aligned.  The library receives raw pointers; after the call, results()
    makes every frame's views, the parameter arrays trimmed to the frame's
    superpixel count, with one split of the buffer."""

    def __init__(self, frames, device):
        # frames: list of (cap, H, W)
        self.frames = frames
        self.fields = []  # per frame: [(offset, aligned bytes)] per field, then spix
        total = 0
        for cap, H, W in frames:
            f = []
            for _, dt, w in _lib.PARAM_FIELDS:
                nb = cap * w * _ITEMSIZE[dt]
                ab = (nb + _ALIGN - 1) // _ALIGN * _ALIGN
                f.append((total, ab))
                total += ab
            ab = (4 * H * W + _ALIGN - 1) // _ALIGN * _ALIGN
            f.append((total, ab))
            total += ab
            self.fields.append(f)
        self.total = total
        self.buf = torch.empty(max(total, 1), dtype=torch.uint8, device=device)
        self.base = self.buf.data_ptr()

    def spix_ptr(self, i: int) -> int:
        return self.base + self.fields[i][-1][0]

    def params_struct(self, i: int) -> _lib.ParamsStruct:
        kw = {name: self.base + off
              for (name, _, _), (off, _) in zip(_lib.PARAM_FIELDS, self.fields[i])}
        return _lib.ParamsStruct(cap=int(self.frames[i][0]), **kw)

    def results(self, ns):
        """[(spix, params)] with params trimmed to ns[i] slots."""
        sizes = []
        for (cap, H, W), f, n in zip(self.frames, self.fields, ns):
            for (_, dt, w), (_, ab) in zip(_lib.PARAM_FIELDS, f):
                used = n * w * _ITEMSIZE[dt]
                sizes += [used, ab - used]
            used = 4 * H * W
            sizes += [used, f[-1][1] - used]
        parts = torch.split(self.buf[:self.total], sizes)
        out = []
        j = 0
        for i, ((cap, H, W), n) in enumerate(zip(self.frames, ns)):
            p = ParamSet()
            for name, dt, w in _lib.PARAM_FIELDS:
                v = parts[j].view(_DTYPES[dt])
                p[name] = v.view(n, w) if w > 1 else v
                j += 2
            spix = parts[j].view(torch.int32).view(H, W)
            j += 2
            # the struct of these arrays as the input of the stream's next
            # frame (valid while the dict still holds these same tensors)
            ps = self.params_struct(i)
            ps.cap = int(n)
            object.__setattr__(p, "_struct", (ps, tuple(p.values())))
            out.append((spix, p))
        return out


def _cached_struct(p: Params, n: int) -> Optional[_lib.ParamsStruct]:
    """The struct made with a batched call's outputs, if p still holds those
    very tensors (no per-field checks needed), else None."""
    c = getattr(p, "__dict__", {}).get("_struct")
    if c is None:
        return None
    ps, tensors = c
    if ps.cap != n or len(p) != len(tensors):
        return None
    for (name, _, _), t in zip(_lib.PARAM_FIELDS, tensors):
        if p.get(name) is not t:
            return None
    return ps


def params_struct(p: Params, cap: Optional[int] = None) -> _lib.ParamsStruct:
    kw = {}
    n = None
    for name, dt, w in _lib.PARAM_FIELDS:
        t = p[name]
        if t.dtype != _DTYPES[dt]:
            raise RuntimeError(f"params['{name}'] must be {dt}")
        _check_input(t, f"params['{name}']")
        n = t.shape[0] if n is None else min(n, t.shape[0])
        kw[name] = t.data_ptr()
    return _lib.ParamsStruct(cap=int(cap if cap is not None else n), **kw)


def trim(p: Params, n: int) -> Params:
    return ParamSet((k, v[:n]) for k, v in p.items())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bass_frame(img: torch.Tensor, cfg: _lib.Config,
               ctx: Optional[Context] = None) -> Tuple[torch.Tensor, Params]:
    """Single-image BASS on one (H, W, 3) float32 CUDA image."""
    _check_input(img, "img")
    if img.dtype != torch.float32 or img.dim() != 3 or img.shape[2] != 3:
        raise RuntimeError("img must be float32 (H, W, 3)")
    H, W = img.shape[:2]
    if ctx is None:
        ctx = get_context(H, W, cfg, img.device)
    else:
        ctx.set_config(cfg)
    cap = _lib.lib().stsp_bass_capacity(cfg.sp_size, H, W)
    out = alloc_params(cap, img.device)
    spix = torch.empty((H, W), dtype=torch.int32, device=img.device)
    n = ctypes.c_int(0)
    with torch.cuda.device(img.device):
        rc = _lib.lib().stsp_bass(ctx.handle, img.data_ptr(), spix.data_ptr(),
                                  ctypes.byref(params_struct(out, cap)), ctypes.byref(n),
                                  _stream())
    _lib.check(rc, "stsp_bass")
    return spix, trim(out, n.value)


def check_prop_inputs(ctx: Optional["Context"], img: torch.Tensor, flow: torch.Tensor,
                      spix_prev: torch.Tensor) -> None:
    """Device, layout, dtype and shape checks of one propagated frame's inputs
    (the kernels index them with the context's frame size)."""
    for t, name in ((img, "img"), (flow, "flow"), (spix_prev, "spix_prev")):
        _check_input(t, name)
    if img.dtype != torch.float32 or img.dim() != 3 or img.shape[2] != 3:
        raise RuntimeError("img must be float32 (H, W, 3)")
    H, W = img.shape[:2]
    if ctx is not None and (H, W) != (ctx.H, ctx.W):
        raise RuntimeError(f"img is {H}x{W} but its context is {ctx.H}x{ctx.W}")
    if flow.shape != (H, W, 2) or flow.dtype != torch.float32:
        raise RuntimeError("flow must be float32 (H, W, 2)")
    if spix_prev.shape != (H, W) or spix_prev.dtype != torch.int32:
        raise RuntimeError("spix_prev must be int32 (H, W)")


def prop_frame(img: torch.Tensor, flow: torch.Tensor, spix_prev: torch.Tensor,
               params_prev: Params, cfg: _lib.Config,
               ctx: Optional[Context] = None, maps: bool = False):
    """One propagated frame.  ``params_prev['prior_counts']`` and
    ``params_prev['invalid_counts']`` are updated in place, as the reference's
    shift_and_fill does.  Returns (spix, params, adaptive_niters); with
    ``maps`` also the shifted labels (-1 at holes) and the filled labels that
    start the refinement (stsp_prop_ex)."""
    check_prop_inputs(ctx, img, flow, spix_prev)
    H, W = img.shape[:2]
    nprev = int(params_prev["ids"].shape[0])
    if ctx is None:
        ctx = get_context(H, W, cfg, img.device)
    else:
        ctx.set_config(cfg)
    cap = _lib.lib().stsp_prop_capacity(nprev)
    out = alloc_params(cap, img.device)
    spix = torch.empty((H, W), dtype=torch.int32, device=img.device)
    n = ctypes.c_int(0)
    ns = ctypes.c_int(0)
    shifted = filled = None
    if maps:
        shifted = torch.empty((H, W), dtype=torch.int32, device=img.device)
        filled = torch.empty((H, W), dtype=torch.int32, device=img.device)
    with torch.cuda.device(img.device):
        rc = _lib.lib().stsp_prop_ex(ctx.handle, img.data_ptr(), flow.data_ptr(),
                                     spix_prev.data_ptr(),
                                     ctypes.byref(params_struct(params_prev, nprev)), nprev,
                                     spix.data_ptr(), ctypes.byref(params_struct(out, cap)),
                                     ctypes.byref(n), ctypes.byref(ns),
                                     shifted.data_ptr() if maps else None,
                                     filled.data_ptr() if maps else None, _stream())
    _lib.check(rc, "stsp_prop_ex")
    if maps:
        return spix, trim(out, n.value), ns.value, shifted, filled
    return spix, trim(out, n.value), ns.value


def marked_video(vid: torch.Tensor, spix: torch.Tensor, color) -> torch.Tensor:
    _check_input(vid, "vid")
    _check_input(spix, "spix")
    T, H, W = spix.shape
    if vid.shape[:3] != spix.shape or vid.shape[3] != 3:
        raise RuntimeError("vid must be (T, H, W, 3) matching spix (T, H, W)")
    c = torch.as_tensor(color, dtype=torch.float32).flatten().cpu()
    if c.numel() != 3:
        raise RuntimeError("color must have 3 entries")
    cbuf = (ctypes.c_float * 3)(*[float(v) for v in c])
    marked = torch.empty_like(vid)
    with torch.cuda.device(vid.device):
        rc = _lib.lib().stsp_marked_video(vid.data_ptr(), spix.data_ptr(), cbuf, T, H, W,
                                          marked.data_ptr(), _stream())
    _lib.check(rc, "stsp_marked_video")
    return marked


class SweepTuner:
    """Autotuner of the fused sweep's launch shape (tile x sweeps per launch).

    Batches are keyed by their composition: the kind of frame (BASS first
    frames or propagated frames), the total pixel count and the largest
    frame, each rounded to half an octave.  The first batch of a key runs
    with the current shape; after its results are out, one probe
    (stsp_batch_tune_sweeps) times every shape on that batch's frames and
    the cheapest becomes the key's shape for every later batch.  Shapes give
    bit-identical results, so the choice changes speed only.  One tuner may
    be shared by several runners."""

    def __init__(self, nlaunch: int = 4):
        self.nlaunch = int(nlaunch)
        self.chosen = {}   # key -> shape index
        self.probes = {}   # key -> microseconds per sweep of every shape
        self._lock = threading.Lock()

    @staticmethod
    def key(kind: str, ctxs) -> tuple:
        px = [c.H * c.W for c in ctxs]
        return (kind, int(round(2 * math.log2(sum(px)))), int(round(2 * math.log2(max(px)))))

    def shape_for(self, key):
        with self._lock:
            return self.chosen.get(key)

    def probe(self, runner, key, stream) -> int:
        us = (ctypes.c_float * _lib.STSP_SWEEP_SHAPES)()
        best = ctypes.c_int()
        _lib.check(_lib.lib().stsp_batch_tune_sweeps(runner.handle, self.nlaunch,
                                                     ctypes.c_void_p(stream.cuda_stream), us,
                                                     ctypes.byref(best)),
                   "stsp_batch_tune_sweeps")
        with self._lock:
            self.chosen[key] = int(best.value)
            self.probes[key] = [round(float(x), 3) for x in us]
        return int(best.value)

    def report(self) -> list:
        with self._lock:
            return [{"kind": k[0], "total_px_log2x2": k[1], "max_px_log2x2": k[2],
                     "shape": self.chosen[k], "sweep": list(SWEEP_SHAPES[self.chosen[k]]),
                     "us_per_sweep": self.probes.get(k)} for k in sorted(self.chosen)]


class BatchRunner:
    """Refines several frames (one per Context) in shared launches
    (stsp_prop_batch, stsp_bass_batch), optionally replaying a captured
    hipGraph and picking the sweep launch shape with a SweepTuner."""

    def __init__(self, max_frames: int = 16, use_graph: bool = False,
                 tuner: Optional["SweepTuner"] = None, defer_probe: bool = False):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().stsp_batch_create(int(max_frames), ctypes.byref(h)),
                   "stsp_batch_create")
        self.handle = h
        self.max_frames = int(max_frames)
        self.tuner = tuner
        # with defer_probe the caller runs a batch's pending probe itself
        # (probe_pending), e.g. after recording the batch's completion
        self.defer_probe = bool(defer_probe)
        self._pending = None
        _lib.check(_lib.lib().stsp_batch_use_graph(h, int(use_graph)), "stsp_batch_use_graph")

    def _tuned(self, kind, ctxs):
        """The tuner's key for this batch; applies its shape if it has one."""
        if self.tuner is None:
            return None
        key = SweepTuner.key(kind, ctxs)
        shape = self.tuner.shape_for(key)
        if shape is not None:
            for c in ctxs:
                c.set_tuning(shape)
            return None
        return key

    def _probe(self, key, ctxs, s):
        if key is None:
            return
        self._pending = (key, ctxs, s)
        if not self.defer_probe:
            self.probe_pending()

    def probe_pending(self) -> None:
        """Runs the tuner's probe over the last batch if it asked for one."""
        if self._pending is None:
            return
        key, ctxs, s = self._pending
        self._pending = None
        shape = self.tuner.probe(self, key, s)
        for c in ctxs:
            c.set_tuning(shape)

    def close(self) -> None:
        if self.handle:
            _lib.lib().stsp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def time_sweeps(self, nlaunch: int, stream=None) -> float:
        """Milliseconds of `nlaunch` back-to-back launches of the fused
        boundary sweep over the frames of the last batch (HIP events on the
        launch stream; stsp_batch_time_sweeps).  For benchmarks."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        ms = ctypes.c_float()
        _lib.check(_lib.lib().stsp_batch_time_sweeps(self.handle, int(nlaunch),
                                                     ctypes.c_void_p(s.cuda_stream),
                                                     ctypes.byref(ms)),
                   "stsp_batch_time_sweeps")
        return float(ms.value)

    def bass(self, items, stream=None):
        """items: list of (ctx, img).  First frames (BASS) of several images or
        streams in shared launches.  Returns a list of (spix, params)."""
        n = len(items)
        if n == 0:
            return []
        if n > self.max_frames:
            raise RuntimeError(f"batch of {n} exceeds max_frames={self.max_frames}")
        VP = ctypes.c_void_p
        ctxs, imgs, souts = (VP * n)(), (VP * n)(), (VP * n)()
        outs = (_lib.ParamsStruct * n)()
        nout = (ctypes.c_int * n)()
        frames = []
        for i, (ctx, img) in enumerate(items):
            _check_input(img, "img")
            if img.dtype != torch.float32 or img.dim() != 3 or img.shape[2] != 3:
                raise RuntimeError("img must be float32 (H, W, 3)")
            if tuple(img.shape[:2]) != (ctx.H, ctx.W):
                raise RuntimeError("img size does not match its context")
            cap = _lib.lib().stsp_bass_capacity(ctx.cfg.sp_size, ctx.H, ctx.W)
            frames.append((cap, ctx.H, ctx.W))
            ctxs[i] = ctx.handle.value
            imgs[i] = img.data_ptr()
        ob = _OutBatch(frames, items[0][1].device)
        for i in range(n):
            souts[i] = ob.spix_ptr(i)
            outs[i] = ob.params_struct(i)
        s = stream if stream is not None else torch.cuda.current_stream()
        key = self._tuned("bass", [it[0] for it in items])
        rc = _lib.lib().stsp_bass_batch(self.handle, ctxs, n, imgs, souts, outs, nout,
                                        ctypes.c_void_p(s.cuda_stream))
        _lib.check(rc, "stsp_bass_batch")
        self._probe(key, [it[0] for it in items], s)
        return ob.results([nout[i] for i in range(n)])

    def prop(self, items, stream=None):
        """items: list of (ctx, img, flow, spix_prev, params_prev).  Returns a
        list of (spix, params, adaptive_niters) in the same order."""
        n = len(items)
        if n == 0:
            return []
        if n > self.max_frames:
            raise RuntimeError(f"batch of {n} exceeds max_frames={self.max_frames}")
        VP = ctypes.c_void_p
        ctxs = (VP * n)()
        imgs = (VP * n)()
        flows = (VP * n)()
        sprev = (VP * n)()
        souts = (VP * n)()
        prevs = (_lib.ParamsStruct * n)()
        outs = (_lib.ParamsStruct * n)()
        nprev = (ctypes.c_int * n)()
        nout = (ctypes.c_int * n)()
        niters = (ctypes.c_int * n)()
        frames = []
        for i, (ctx, img, flow, spix_prev, params_prev) in enumerate(items):
            check_prop_inputs(ctx, img, flow, spix_prev)
            np_ = int(params_prev["ids"].shape[0])
            cap = _lib.lib().stsp_prop_capacity(np_)
            frames.append((cap, ctx.H, ctx.W))
            ctxs[i] = ctx.handle.value
            imgs[i] = img.data_ptr()
            flows[i] = flow.data_ptr()
            sprev[i] = spix_prev.data_ptr()
            prevs[i] = _cached_struct(params_prev, np_) or params_struct(params_prev, np_)
            nprev[i] = np_
        ob = _OutBatch(frames, items[0][1].device)
        for i in range(n):
            souts[i] = ob.spix_ptr(i)
            outs[i] = ob.params_struct(i)
        s = stream if stream is not None else torch.cuda.current_stream()
        key = self._tuned("prop", [it[0] for it in items])
        rc = _lib.lib().stsp_prop_batch(self.handle, ctxs, n, imgs, flows, sprev, prevs, nprev,
                                        souts, outs, nout, niters, ctypes.c_void_p(s.cuda_stream))
        _lib.check(rc, "stsp_prop_batch")
        self._probe(key, [it[0] for it in items], s)
        res = ob.results([nout[i] for i in range(n)])
        return [(res[i][0], res[i][1], niters[i]) for i in range(n)]
