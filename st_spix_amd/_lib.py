"""ctypes binding of libstspix.so (the C ABI declared in include/stspix.h).

The library is built in-tree by ``st_spix_amd/csrc/build.sh`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing, importing the compute entry points raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libstspix.so")

STATUS = {0: "ok", -1: "invalid argument", -2: "output capacity too small",
          -3: "hole filling stopped shrinking", -4: "HIP runtime error",
          -5: "invalid stream state", -6: "frame exceeds the static scratch buffer"}


class StspError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: {STATUS.get(code, 'unknown')} ({code})")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("niters", ctypes.c_int), ("niters_seg", ctypes.c_int),
                ("sm_start", ctypes.c_int), ("sp_size", ctypes.c_int),
                ("sm_period", ctypes.c_int), ("prop_niters", ctypes.c_int),
                ("sigma2_app", ctypes.c_float), ("potts", ctypes.c_float),
                ("alpha_hastings", ctypes.c_float), ("split_alpha", ctypes.c_float),
                ("merge_alpha", ctypes.c_float), ("thresh_relabel", ctypes.c_float),
                ("thresh_new", ctypes.c_float), ("target_nspix", ctypes.c_int)]


PARAM_FIELDS = [
    ("mu_app", "float32", 3),
    ("prior_mu_app", "float32", 3),
    ("mu_shape", "float64", 2),
    ("sigma_shape", "float64", 3),
    ("logdet_sigma_shape", "float32", 1),
    ("prior_mu_shape", "float64", 2),
    ("prior_sigma_shape", "float64", 3),
    ("sample_sigma_shape", "float64", 3),
    ("counts", "int32", 1),
    ("sm_counts", "int32", 1),
    ("prior_counts", "float32", 1),
    ("ids", "int32", 1),
    ("invalid_counts", "float32", 1),
]


class ParamsStruct(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n, _, _ in PARAM_FIELDS] + [("cap", ctypes.c_int)]


_lib = None

# every symbol include/stspix.h declares
# number of fused-sweep launch shapes (stsp_ctx_set_tuning, stsp_batch_tune_sweeps)
STSP_SWEEP_SHAPES = 6

EXPORTS = ["stsp_default_config", "stsp_ctx_create", "stsp_ctx_destroy",
           "stsp_ctx_set_config", "stsp_bass_capacity", "stsp_prop_capacity", "stsp_bass",
           "stsp_prop", "stsp_marked_video", "stsp_status_string", "stsp_ctx_enable_timing",
           "stsp_ctx_kernel_time", "stsp_ctx_set_tuning", "stsp_batch_create",
           "stsp_batch_destroy", "stsp_batch_use_graph", "stsp_batch_time_sweeps", "stsp_batch_tune_sweeps", "stsp_shift_labels", "stsp_fill_missing", "stsp_ctx_set_shape_marginal", "stsp_prop_batch", "stsp_bass_batch",
           "stsp_rescale_image", "stsp_rgb2lab", "stsp_lab2rgb", "stsp_ctx_set_split_search",
           "stsp_ctx_split_search_counts", "stsp_bass_noalloc", "stsp_bass_noalloc_scratch_bytes",
           "stsp_prop_ex", "stsp_ctx_device_bytes"]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with st_spix_amd/csrc/build.sh "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, pi = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)
        L.stsp_default_config.argtypes = [ctypes.POINTER(Config)]
        L.stsp_default_config.restype = None
        L.stsp_ctx_create.argtypes = [ctypes.POINTER(Config), i, i, ctypes.POINTER(vp)]
        L.stsp_ctx_destroy.argtypes = [vp]
        L.stsp_ctx_destroy.restype = None
        L.stsp_ctx_set_config.argtypes = [vp, ctypes.POINTER(Config)]
        L.stsp_bass_capacity.argtypes = [i, i, i]
        L.stsp_prop_capacity.argtypes = [i]
        L.stsp_bass.argtypes = [vp, vp, vp, ctypes.POINTER(ParamsStruct), pi, vp]
        L.stsp_prop.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ParamsStruct), i, vp,
                                ctypes.POINTER(ParamsStruct), pi, pi, vp]
        L.stsp_ctx_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
        L.stsp_prop_ex.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ParamsStruct), i, vp,
                                   ctypes.POINTER(ParamsStruct), pi, pi, vp, vp, vp]
        L.stsp_marked_video.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_float), i, i, i, vp, vp]
        L.stsp_ctx_enable_timing.argtypes = [vp, i]
        L.stsp_ctx_kernel_time.argtypes = [vp, i, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_longlong)]
        L.stsp_ctx_set_tuning.argtypes = [vp, i]
        L.stsp_ctx_set_split_search.argtypes = [vp, i, i, i]
        L.stsp_ctx_set_shape_marginal.argtypes = [vp, i]
        L.stsp_ctx_split_search_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
        L.stsp_bass_noalloc.argtypes = [ctypes.POINTER(Config), i, i, vp, vp,
                                        ctypes.POINTER(ParamsStruct), pi, vp]
        L.stsp_bass_noalloc_scratch_bytes.argtypes = [ctypes.POINTER(Config), i, i,
                                                      ctypes.POINTER(ctypes.c_size_t)]
        L.stsp_batch_create.argtypes = [i, ctypes.POINTER(vp)]
        L.stsp_batch_destroy.argtypes = [vp]
        L.stsp_batch_destroy.restype = None
        L.stsp_batch_use_graph.argtypes = [vp, i]
        L.stsp_batch_time_sweeps.argtypes = [vp, i, vp, ctypes.POINTER(ctypes.c_float)]
        L.stsp_batch_tune_sweeps.argtypes = [vp, i, vp, ctypes.POINTER(ctypes.c_float),
                                             ctypes.POINTER(i)]
        L.stsp_shift_labels.argtypes = [vp, vp, i, i, i, i, vp, vp, vp]
        L.stsp_fill_missing.argtypes = [vp, vp, i, vp, i, i, i, i, pi, pi, vp]
        L.stsp_prop_batch.argtypes = [vp, ctypes.POINTER(vp), i, ctypes.POINTER(vp),
                                      ctypes.POINTER(vp), ctypes.POINTER(vp),
                                      ctypes.POINTER(ParamsStruct), pi, ctypes.POINTER(vp),
                                      ctypes.POINTER(ParamsStruct), pi, pi, vp]
        L.stsp_bass_batch.argtypes = [vp, ctypes.POINTER(vp), i, ctypes.POINTER(vp),
                                      ctypes.POINTER(vp), ctypes.POINTER(ParamsStruct), pi, vp]
        L.stsp_rescale_image.argtypes = [vp, vp, i, i, ctypes.c_float, vp]
        L.stsp_rgb2lab.argtypes = [vp, vp, i, i, vp]
        L.stsp_lab2rgb.argtypes = [vp, vp, i, i, vp]
        L.stsp_status_string.argtypes = [i]
        L.stsp_status_string.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != 0:
        raise StspError(code, where)
