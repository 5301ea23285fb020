"""stsp_bass_noalloc (single-image BASS whose workspace is a static device
buffer) against stsp_bass, through the C ABI: bit-identical labels and
parameters on consecutive calls of decreasing size, no device memory taken
by the calls, and a clear error beyond the buffer's size."""
import ctypes

import numpy as np
import pytest
import torch

from st_spix_amd import _lib, core, synth

pytestmark = pytest.mark.gpu

FIELDS = [n for n, _, _ in _lib.PARAM_FIELDS]


def noalloc_bass(img, cfg):
    L = _lib.lib()
    H, W = img.shape[:2]
    cap = L.stsp_bass_capacity(cfg.sp_size, H, W)
    out = core.alloc_params(cap, img.device)
    spix = torch.empty((H, W), dtype=torch.int32, device=img.device)
    n = ctypes.c_int(0)
    rc = L.stsp_bass_noalloc(ctypes.byref(cfg), H, W, img.data_ptr(), spix.data_ptr(),
                             ctypes.byref(core.params_struct(out, cap)), ctypes.byref(n),
                             core._stream())
    return rc, spix, core.trim(out, n.value)


def assert_same(a, pa, b, pb, what):
    assert torch.equal(a, b), f"{what}: labels differ"
    for f in FIELDS:
        x, y = pa[f].cpu().numpy(), pb[f].cpu().numpy()
        assert x.shape == y.shape and np.array_equal(x, y, equal_nan=True), f"{what}: {f} differs"


def test_noalloc_matches_bass_two_calls():
    cases = [((160, 224, 14, 7), dict(niters=40, sp_size=12, potts=10.0)),   # larger first
             ((48, 64, 6, 3), dict(niters=24, sp_size=8, potts=10.0))]      # then smaller
    for i, ((H, W, nreg, seed), kw) in enumerate(cases):
        img = torch.from_numpy(synth.voronoi_image(H, W, nreg, seed=seed)).cuda()
        cfg = core.make_config(**kw)
        ref, pref = core.bass_frame(img, cfg)
        rc, got, pgot = noalloc_bass(img, cfg)
        torch.cuda.synchronize()
        _lib.check(rc, "stsp_bass_noalloc")
        assert_same(got, pgot, ref, pref, f"call {i} ({H}x{W})")


def test_noalloc_takes_no_device_memory():
    img = torch.from_numpy(synth.voronoi_image(96, 128, 9, seed=2)).cuda()
    small = img[:40, :50].contiguous()
    cfg = core.make_config(niters=12, sp_size=10, potts=10.0)
    rc, _, _ = noalloc_bass(img, cfg)  # warm: the runtime's own first-use set-up
    torch.cuda.synchronize()
    _lib.check(rc, "stsp_bass_noalloc")
    L = _lib.lib()
    outs = []
    for im in (img, small):  # outputs allocated before the measured calls
        H, W = im.shape[:2]
        cap = L.stsp_bass_capacity(cfg.sp_size, H, W)
        outs.append((core.alloc_params(cap, im.device),
                     torch.empty((H, W), dtype=torch.int32, device=im.device), cap))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for im, (out, spix, cap) in zip((img, small), outs):
        n = ctypes.c_int(0)
        rc = L.stsp_bass_noalloc(ctypes.byref(cfg), im.shape[0], im.shape[1], im.data_ptr(),
                                 spix.data_ptr(), ctypes.byref(core.params_struct(out, cap)),
                                 ctypes.byref(n), core._stream())
        _lib.check(rc, "stsp_bass_noalloc")
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_rejects_frames_beyond_the_buffer():
    L = _lib.lib()
    cfg = core.make_config(niters=4, sp_size=25)
    need = ctypes.c_size_t(0)
    _lib.check(L.stsp_bass_noalloc_scratch_bytes(ctypes.byref(cfg), 1080, 1920,
                                                 ctypes.byref(need)), "scratch bytes")
    assert need.value > 32 << 20
    img = torch.zeros((1080, 1920, 3), dtype=torch.float32, device="cuda")
    rc, _, _ = noalloc_bass(img, cfg)
    assert rc == -6
    assert "scratch" in L.stsp_status_string(rc).decode()
