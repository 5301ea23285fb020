"""Host-side serving behaviour that needs no GPU: admission against the HBM
budget and request bookkeeping."""
import math

import pytest

from st_spix_amd import serve


def test_admission_budget():
    a = serve.Admission(1000)
    a.reserve(600)
    with pytest.raises(serve.AdmissionError):
        a.reserve(500)
    a.release(600)
    a.reserve(1000)
    assert a.used == 1000


def test_stream_bytes_monotone():
    small = serve.estimate_stream_bytes(360, 640, 25)
    big = serve.estimate_stream_bytes(1080, 1920, 25)
    assert 0 < small < big
    # 16 mixed-resolution streams fit the 64 GB budget with room to spare
    mix = sum(serve.estimate_stream_bytes(h, w, 25) for h, w in
              [(360, 640), (480, 854), (720, 1280), (1080, 1920)]) * 4
    assert mix < 64 * (1 << 30) / 4


def test_latency_percentiles():
    reqs = [serve.Request(0, None, None, arrival=0.0, done=0.001 * (i + 1)) for i in range(100)]
    p = serve.latency_percentiles(reqs)
    assert math.isclose(p[50], 0.050) and math.isclose(p[99], 0.099)


def test_input_checks_raise_without_gpu():
    import torch

    import st_spix_amd
    vid = torch.zeros((2, 8, 8, 3))
    with pytest.raises(RuntimeError, match="CUDA"):
        st_spix_amd.bist_forward(vid, torch.zeros((1, 8, 8, 2)), 4, 5, 1.0, 0.01, -0.7, 0.0,
                                 True, False)
