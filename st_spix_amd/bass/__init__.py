"""Per-frame single-image BASS entry point (the ``prop_cuda.bass`` call shape
used by the reference's st_spix/prop/__init__.py:64-66 and
st_spix/models/bass.py:56-58)."""
from __future__ import annotations

import math

import torch

from .. import core


def bass(img: torch.Tensor, niters: int = 500, niters_seg: int = 4, sm_start: int = 0,
         sp_size: int = 25, sigma2_app: float = 8.1e-5, sigma2_size: float = 0.0,
         potts: float = 0.5, alpha_hastings: float = math.log(0.5),
         split_alpha: float = 0.0, **extra):
    """img: (1, H, W, 3) or (H, W, 3) float32 CUDA.  Returns (spix, params):
    spix int32 with the input's leading shape, params a dict of the
    SuperpixelParams fields (structs.h:13-48).  ``sigma2_size`` is accepted
    and unused, as in the reference (pyapi.cu:60)."""
    batched = img.dim() == 4
    if batched:
        if img.shape[0] != 1:
            raise RuntimeError("bass expects a single image (batch size 1)")
        img = img[0]
    cfg = core.make_config(niters=niters, niters_seg=niters_seg, sm_start=sm_start,
                           sp_size=sp_size, sigma2_app=sigma2_app, potts=potts,
                           alpha=alpha_hastings, split_alpha=split_alpha, **extra)
    spix, params = core.bass_frame(img.contiguous(), cfg)
    return (spix[None] if batched else spix), params
