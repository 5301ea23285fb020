"""Colour preprocessing of the reference's command-line path (demo.cu:248-250):
uint8 rescale and the sRGB <-> Lab conversions of rgb2lab.cu, on CUDA tensors
through the C ABI.  ``rgb2lab`` reads channels in B,G,R order (OpenCV) and
returns (L/-100, a/100, b/100); ``lab2rgb`` returns R,G,B in [0, 1]."""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .core import _check_input, _stream


def _npix(t: torch.Tensor):
    if t.dim() < 3 or t.shape[-1] != 3:
        raise RuntimeError("expected (..., H, W, 3)")
    H, W = t.shape[-3], t.shape[-2]
    n = t.numel() // (3 * H * W)
    return n, H * W


def rescale_image(img_u8: torch.Tensor, scale: float = 255.0) -> torch.Tensor:
    """rescale_image (rgb2lab.cu:19-32): float32 copy divided by scale."""
    _check_input(img_u8, "img")
    if img_u8.dtype != torch.uint8:
        raise RuntimeError("img must be uint8")
    n, npix = _npix(img_u8)
    out = torch.empty(img_u8.shape, dtype=torch.float32, device=img_u8.device)
    with torch.cuda.device(img_u8.device):
        _lib.check(_lib.lib().stsp_rescale_image(img_u8.data_ptr(), out.data_ptr(), n, npix,
                                                 ctypes.c_float(scale), _stream()),
                   "stsp_rescale_image")
    return out


def rgb2lab(img_bgr: torch.Tensor) -> torch.Tensor:
    """rgb2lab (rgb2lab.cu:35-117)."""
    _check_input(img_bgr, "img")
    if img_bgr.dtype != torch.float32:
        raise RuntimeError("img must be float32")
    n, npix = _npix(img_bgr)
    out = torch.zeros_like(img_bgr)
    with torch.cuda.device(img_bgr.device):
        _lib.check(_lib.lib().stsp_rgb2lab(img_bgr.data_ptr(), out.data_ptr(), n, npix,
                                           _stream()), "stsp_rgb2lab")
    return out


def lab2rgb(img_lab: torch.Tensor) -> torch.Tensor:
    """lab2rgb (rgb2lab.cu:42-162)."""
    _check_input(img_lab, "img")
    if img_lab.dtype != torch.float32:
        raise RuntimeError("img must be float32")
    n, npix = _npix(img_lab)
    out = torch.zeros_like(img_lab)
    with torch.cuda.device(img_lab.device):
        _lib.check(_lib.lib().stsp_lab2rgb(img_lab.data_ptr(), out.data_ptr(), n, npix,
                                           _stream()), "stsp_lab2rgb")
    return out
