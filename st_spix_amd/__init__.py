"""MI355X-native streaming BASS/BIST superpixels.

Drop-in for the reference's ``bist_cuda`` module (bist_forward,
get_marked_video; st_spix/csrc/pybind_bist.cpp:16-17) and for the per-frame
``st_spix.bass`` / ``st_spix.prop`` entry points, computed by hand-written
HIP kernels in ``_lib/libstspix.so``.  ``st_spix_amd.bist_cuda`` is the thin
torch C++ extension with the reference module's two functions; the
``bist_forward`` / ``get_marked_video`` below call it, and take the
ctypes binding of the same library (core) only for keyword overrides of the
constants the reference hard-codes.
"""
from __future__ import annotations

import math

import torch

from . import core
from .core import Context, alloc_params, make_config  # noqa: F401

__all__ = ["bist_forward", "get_marked_video", "make_config", "Context"]


def bist_forward(vid: torch.Tensor, flows: torch.Tensor, niters: int, sp_size: int,
                 potts: float, sigma2_app: float, alpha: float, split_alpha: float,
                 video_mode: bool, rgb2lab: bool = False, **extra) -> torch.Tensor:
    """Superpixels of a video (reference pyapi.cu:200-215 / main_loop 46-195).

    vid: (T, H, W, 3) float32 CUDA; flows: (>=T-1, H, W, 2) float32 CUDA with
    flows[t-1] the flow into frame t.  Returns int32 labels (T, H, W).
    ``rgb2lab`` is accepted and ignored, as in the reference (pyapi.cu:97-101).
    Extra keyword arguments (niters_seg, sm_period, prop_niters, ...) override
    constants the reference hard-codes.
    """
    if not extra:
        return _ext().bist_forward(vid, flows, int(niters), int(sp_size), float(potts),
                                   float(sigma2_app), float(alpha), float(split_alpha),
                                   bool(video_mode), bool(rgb2lab))
    core._check_input(vid, "vid")
    core._check_input(flows, "flows")
    T, H, W = vid.shape[:3]
    cfg = make_config(niters=niters, sp_size=sp_size, potts=potts, sigma2_app=sigma2_app,
                      alpha=alpha, split_alpha=split_alpha, **extra)
    out = torch.zeros((T, H, W), dtype=torch.int32, device=vid.device)
    spix_prev = params_prev = None
    for t in range(T):
        if t == 0 or not video_mode:
            spix, params = core.bass_frame(vid[t], cfg)
        else:
            spix, params, _ = core.prop_frame(vid[t], flows[t - 1], spix_prev, params_prev, cfg)
        out[t] = spix
        spix_prev, params_prev = spix, params
    return out


def get_marked_video(vid: torch.Tensor, spix: torch.Tensor, color) -> torch.Tensor:
    """Border overlay (reference pyapi.cu:260-303)."""
    if not torch.is_tensor(color):
        color = torch.tensor(color, dtype=torch.float32)
    return _ext().get_marked_video(vid, spix, color.contiguous())


def _ext():
    """The torch extension (bist_cuda*.so, built by csrc/build_ext.sh); a
    missing build is an error, not a fallback."""
    try:
        from . import bist_cuda
    except ImportError as e:
        raise ImportError("st_spix_amd.bist_cuda is not built: run "
                          "st_spix_amd/csrc/build_ext.sh (or __graft_entry__.build())") from e
    return bist_cuda
