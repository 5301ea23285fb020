"""The per-step label ops of the reference's prop_cuda extension, as used by
st_spix/prop/__init__.py (shift_labels at 366, fill_missing at 463), on
libstspix.so (stsp_shift_labels, stsp_fill_missing in include/stspix.h).

Same argument shapes, dtypes and returns as the reference bindings
(csrc/prop/shift_labels.cu:81-126, csrc/prop/fill_missing.cu:317-354);
inputs must be contiguous CUDA tensors, as the reference's CHECK_INPUT
requires, and a wrong shape or dtype raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import warnings

import torch

from . import _lib


def _check(t: torch.Tensor, name: str, dtype, dim: int) -> None:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}")
    if t.dim() != dim:
        raise RuntimeError(f"{name} must have {dim} dimensions")


def shift_labels(spix: torch.Tensor, flow: torch.Tensor):
    """spix (B, H, W) int32 and flow (B, nspix, 2) int32 ((dx, dy) per
    label) -> (shifted_spix, counts), both (B, H, W) int32: every pixel
    moves by its label's flow; a destination counts its arrivals and keeps
    the largest arriving label (-1 where none arrives)."""
    _check(spix, "spix", torch.int32, 3)
    _check(flow, "flow", torch.int32, 3)
    B, H, W = spix.shape
    if flow.shape[0] != B or flow.shape[2] != 2:
        raise RuntimeError("flow must be (nbatch, nspix, 2)")
    shifted = torch.empty_like(spix)
    counts = torch.empty_like(spix)
    s = torch.cuda.current_stream(spix.device)
    with torch.cuda.device(spix.device):
        rc = _lib.lib().stsp_shift_labels(spix.data_ptr(), flow.data_ptr(), B, H, W,
                                          int(flow.shape[1]), shifted.data_ptr(),
                                          counts.data_ptr(), ctypes.c_void_p(s.cuda_stream))
    _lib.check(rc, "stsp_shift_labels")
    return shifted, counts


def fill_missing(spix: torch.Tensor, centers: torch.Tensor, missing: torch.Tensor,
                 break_iter: int) -> torch.Tensor:
    """spix (1, H, W) int32 with -1 holes, centers (nspix, 2) float32 (x, y
    of each label), missing (1, nmissing) int32 flat indices of the holes
    -> a filled copy of spix.  Holes are filled from their 4-neighbours by
    the nearest centre, in rounds of four checkerboard passes, at most
    break_iter rounds (<= 0: until done).  Warns, as the reference prints,
    when listed holes remain."""
    _check(spix, "spix", torch.int32, 3)
    _check(centers, "centers", torch.float32, 2)
    _check(missing, "missing", torch.int32, 2)
    if spix.shape[0] != 1:
        raise RuntimeError("fill_missing takes one image (nbatch == 1)")
    if centers.shape[1] != 2:
        raise RuntimeError("centers must be (nspix, 2)")
    _, H, W = spix.shape
    filled = spix.clone()
    rounds, neg = ctypes.c_int(), ctypes.c_int()
    s = torch.cuda.current_stream(spix.device)
    with torch.cuda.device(spix.device):
        rc = _lib.lib().stsp_fill_missing(filled.data_ptr(), centers.data_ptr(),
                                          int(centers.shape[0]), missing.data_ptr(),
                                          int(missing.shape[1]), H, W, int(break_iter),
                                          ctypes.byref(rounds), ctypes.byref(neg),
                                          ctypes.c_void_p(s.cuda_stream))
    _lib.check(rc, "stsp_fill_missing")
    if neg.value > 0:
        warnings.warn("negative spix exist.")
    return filled
