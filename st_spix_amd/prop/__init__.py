"""Streaming propagation entry points with the call shapes of the
reference's st_spix/prop/__init__.py (indepent_bass 38-85, stream_bass
87-282, run_prop 284-656), computed by libstspix.so.

The reference's bodies run the legacy prop_cuda refinement; these run the
BIST frame of st_spix/csrc/bist (shift_and_fill + run_prop,
pyapi.cu:128-170) behind the same arguments and return values.  Argument
order, defaults and the per-call overrides the reference makes inside the
functions (``niters = sp_size`` for the first frame, ``sm_start = 0``,
``alpha_hastings = -1.3`` for propagated frames in stream_bass) are kept.
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import torch

from .. import core
from ..bass import bass as _bass

Params = core.ParamSet


def _cfg(niters, niters_seg, sm_start, sp_size, sigma2_app, potts, alpha_hastings, **extra):
    return core.make_config(niters=niters, niters_seg=niters_seg, sm_start=sm_start,
                            sp_size=sp_size, sigma2_app=sigma2_app, potts=potts,
                            alpha=alpha_hastings, **extra)


def run_prop(img: torch.Tensor, flow: torch.Tensor, spix_tm1: torch.Tensor, params_tm1,
             niters: int, niters_seg: int, sp_size: int, sigma2_app: float,
             sigma2_size: float, potts: float, alpha_hastings: float, sm_start: int = 0,
             **extra) -> Tuple[torch.Tensor, Params, torch.Tensor, torch.Tensor]:
    """One propagated frame (reference argument order,
    st_spix/prop/__init__.py:284-285).

    img (1, H, W, 3) or (H, W, 3), flow (1, H, W, 2) or (H, W, 2), spix_tm1
    (1, H, W) or (H, W), params_tm1 the previous frame's parameters.
    ``niters`` > 0 fixes the refinement iterations; 0 keeps the BIST
    adaptive 12/8/4 schedule (pyapi.cu:141-149).  ``sigma2_size`` is accepted
    and unused, as in the reference.  params_tm1's prior and invalid counts
    are updated in place, as shift_and_fill does.

    Returns (spix_t, params_t, pre_fill, post_fill) as the reference does
    (:656): pre_fill is the flow-shifted map with -1 at holes and
    collisions, post_fill the filled map that starts the refinement; the
    maps keep the input's leading batch dimension.
    """
    batched = img.dim() == 4
    if batched:
        if img.shape[0] != 1:
            raise RuntimeError("run_prop expects a single frame (batch size 1)")
        img, flow, spix_tm1 = img[0], flow[0], spix_tm1[0]
    cfg = _cfg(niters, niters_seg, sm_start, sp_size, sigma2_app, potts, alpha_hastings,
               prop_niters=max(0, int(niters)), **extra)
    spix, params, _, pre_fill, post_fill = core.prop_frame(
        img.contiguous(), flow.contiguous(), spix_tm1.contiguous(), params_tm1, cfg, maps=True)
    if batched:
        return spix[None], params, pre_fill[None], post_fill[None]
    return spix, params, pre_fill, post_fill


def stream_bass(vid: torch.Tensor, flow: Optional[torch.Tensor] = None, niters: int = 30,
                niters_seg: int = 4, sp_size: int = 25, sigma2_app: float = 0.01,
                sigma2_size: float = 0.1, alpha_hastings: float = 0.1, potts: float = 10.,
                sm_start: int = 0, rgb2lab: bool = True, **extra):
    """Whole video: BASS on frame 0, propagation afterwards
    (st_spix/prop/__init__.py:87-282).  vid (T, H, W, 3), flow (T-1, H, W, 2).

    As the reference does, frame 0 runs ``sp_size`` iterations with
    sm_start 0 and the propagated frames use ``alpha_hastings = -1.3``
    (:139-151); ``rgb2lab`` is accepted and unused (its conversion is
    commented out there).  Returns (spix (T, H, W), [params_t],
    pre_fill (T-1, H, W), post_fill (T-1, H, W)).
    """
    if flow is None:
        raise RuntimeError("stream_bass needs the optical flow (the reference's RAFT "
                           "fallback is not available)")
    T = vid.shape[0]
    niters = sp_size
    sm_start = 0
    spix, params = _bass(vid[0], niters, niters_seg, sm_start, sp_size, sigma2_app,
                         sigma2_size, potts, alpha_hastings, **extra)
    alpha_hastings = -1.3
    spix_all: List[torch.Tensor] = [spix]
    plist: List[Params] = [params]
    pre, post = [], []
    for ix in range(T - 1):
        spix, params, pre_t, post_t = run_prop(vid[ix + 1], flow[ix], spix_all[-1], plist[-1],
                                               niters, niters_seg, sp_size, sigma2_app,
                                               sigma2_size, potts, alpha_hastings, sm_start,
                                               **extra)
        spix_all.append(spix)
        plist.append(params)
        pre.append(pre_t)
        post.append(post_t)
    H, W = vid.shape[1:3]
    empty = torch.empty((0, H, W), dtype=torch.int32, device=vid.device)
    return (torch.stack(spix_all), plist, torch.stack(pre) if pre else empty,
            torch.stack(post) if post else empty)


def indepent_bass(vid: torch.Tensor, niters: int = 30, niters_seg: int = 5, sp_size: int = 25,
                  sigma2_app: float = 0.1, sigma2_size: float = 0.1,
                  alpha_hastings: float = 0.01, potts: float = 10., sm_start: int = 0,
                  rgb2lab: bool = True, **extra) -> torch.Tensor:
    """Independent BASS per frame (st_spix/prop/__init__.py:38-85), with the
    reference's overrides ``niters, niters_seg = sp_size, 4`` and
    ``sm_start = 0``.  Returns (T, H, W) int32."""
    niters, niters_seg, sm_start = sp_size, 4, 0
    return torch.stack([_bass(vid[t], niters, niters_seg, sm_start, sp_size, sigma2_app,
                              sigma2_size, potts, alpha_hastings, **extra)[0]
                        for t in range(vid.shape[0])])
