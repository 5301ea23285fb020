"""The C-ABI library loads and exports every symbol include/stspix.h declares
(no compute calls; runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "stspix.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(stsp_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    from st_spix_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libstspix.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_host_only_entry_points():
    from st_spix_amd import _lib
    L = _lib.lib()
    cfg = _lib.Config()
    L.stsp_default_config(ctypes.byref(cfg))
    assert cfg.niters_seg == 4 and cfg.sm_period == 4
    assert abs(cfg.thresh_relabel - 1e-5) < 1e-12
    assert L.stsp_bass_capacity(25, 321, 481) == 30 * 285
    assert L.stsp_prop_capacity(726) == 7260
    assert L.stsp_status_string(-3) == b"hole filling stopped shrinking"
    h = ctypes.c_void_p()
    assert L.stsp_ctx_create(None, 10, 10, ctypes.byref(h)) == -1
    cfg.sp_size = 0
    assert L.stsp_ctx_create(ctypes.byref(cfg), 10, 10, ctypes.byref(h)) == -1


def test_torch_extension_loads():
    """The thin torch C++ extension (bist_cuda, the reference module's two
    functions) imports, links libstspix.so and checks its inputs on the host
    the way the reference's CHECK_INPUT does."""
    import torch
    from st_spix_amd import bist_cuda
    assert callable(bist_cuda.bist_forward) and callable(bist_cuda.get_marked_video)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        bist_cuda.bist_forward(torch.zeros(1, 4, 4, 3), torch.zeros(1, 4, 4, 2), 10, 8, 1.0,
                               0.01, -0.6, 0.0, True)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        bist_cuda.get_marked_video(torch.zeros(1, 4, 4, 3), torch.zeros(1, 4, 4, dtype=torch.int32),
                                   torch.zeros(3))
