"""Streaming propagation entry points mirroring the reference's
st_spix/prop/__init__.py (stream_bass 87-282, run_prop 284-656,
indepent_bass 38-85), computed by libstspix.so."""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Tuple

import torch

from .. import core
from ..bass import bass as _bass

Params = Dict[str, torch.Tensor]


def _cfg(niters, niters_seg, sm_start, sp_size, sigma2_app, potts, alpha_hastings,
         split_alpha=0.0, **extra):
    return core.make_config(niters=niters, niters_seg=niters_seg, sm_start=sm_start,
                            sp_size=sp_size, sigma2_app=sigma2_app, potts=potts,
                            alpha=alpha_hastings, split_alpha=split_alpha, **extra)


def run_prop(img: torch.Tensor, flow: torch.Tensor, spix_tm1: torch.Tensor,
             params_tm1: Params, niters: int = 0, niters_seg: int = 4, sm_start: int = 0,
             sp_size: int = 25, sigma2_app: float = 8.1e-5, sigma2_size: float = 0.0,
             potts: float = 0.5, alpha_hastings: float = math.log(0.5),
             split_alpha: float = 0.0, **extra) -> Tuple[torch.Tensor, Params, int]:
    """One propagated frame: shift spix_tm1 along ``flow`` (H, W, 2), fill,
    refine.  ``niters`` > 0 fixes the refinement iterations; 0 keeps the
    reference's adaptive 12/8/4 schedule.  Returns (spix_t, params_t,
    adaptive_niters); params_tm1's prior/invalid counts are updated in place
    as the reference does."""
    squeeze = img.dim() == 4
    if squeeze:
        img, flow, spix_tm1 = img[0], flow[0], spix_tm1[0]
    cfg = _cfg(niters, niters_seg, sm_start, sp_size, sigma2_app, potts, alpha_hastings,
               split_alpha, prop_niters=niters, **extra)
    spix, params, ns = core.prop_frame(img.contiguous(), flow.contiguous(),
                                       spix_tm1.contiguous(), params_tm1, cfg)
    return (spix[None] if squeeze else spix), params, ns


def stream_bass(vid: torch.Tensor, flow: torch.Tensor, niters: int = 500,
                niters_seg: int = 4, sp_size: int = 25, sigma2_app: float = 8.1e-5,
                sigma2_size: float = 0.0, alpha_hastings: float = math.log(0.5),
                potts: float = 0.5, sm_start: int = 0, rgb2lab: bool = False,
                prop_niters: int = 0, **extra) -> Tuple[torch.Tensor, List[Params]]:
    """Whole video: BASS on frame 0, propagation afterwards.  vid (T, H, W, 3),
    flow (T-1, H, W, 2).  Returns (spix (T, H, W), [params_t])."""
    T = vid.shape[0]
    spix_all, plist = [], []
    spix, params = _bass(vid[0], niters, niters_seg, sm_start, sp_size, sigma2_app,
                         sigma2_size, potts, alpha_hastings, **extra)
    spix_all.append(spix)
    plist.append(params)
    for t in range(1, T):
        spix, params, _ = run_prop(vid[t], flow[t - 1], spix, params, prop_niters,
                                   niters_seg, sm_start, sp_size, sigma2_app, sigma2_size,
                                   potts, alpha_hastings, **extra)
        spix_all.append(spix)
        plist.append(params)
    return torch.stack(spix_all), plist


def indepent_bass(vid: torch.Tensor, niters: int = 500, niters_seg: int = 4,
                  sp_size: int = 25, sigma2_app: float = 8.1e-5, sigma2_size: float = 0.0,
                  alpha_hastings: float = math.log(0.5), potts: float = 0.5,
                  sm_start: int = 0, rgb2lab: bool = False, **extra) -> torch.Tensor:
    """Independent BASS per frame; returns (T, H, W) int32."""
    return torch.stack([_bass(vid[t], niters, niters_seg, sm_start, sp_size, sigma2_app,
                              sigma2_size, potts, alpha_hastings, **extra)[0]
                        for t in range(vid.shape[0])])
