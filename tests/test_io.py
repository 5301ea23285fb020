"""On-disk formats (file_io.cpp, demo.cu save_spix / save_params)."""
import os

import numpy as np
import pytest

from st_spix_amd import io


def test_flo_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(0)
    flow = rng.normal(0, 3, (5, 7, 2)).astype(np.float32)
    p = tmp_path / "00000.flo"
    io.write_flo(str(p), flow)
    raw = p.read_bytes()
    # tag, width, height, then u,v pairs row-major (file_io.cpp:60-90)
    assert np.frombuffer(raw[:4], np.float32)[0] == np.float32(202021.25)
    assert tuple(np.frombuffer(raw[4:12], np.int32)) == (7, 5)
    assert np.frombuffer(raw[12:20], np.float32).tolist() == flow[0, 0].tolist()
    assert len(raw) == 12 + 5 * 7 * 2 * 4
    np.testing.assert_array_equal(io.read_flo(str(p)), flow)


def test_flo_rejects_bad_tag_and_nan(tmp_path):
    p = tmp_path / "bad.flo"
    p.write_bytes(np.array([1.0], np.float32).tobytes() + np.array([2, 2], np.int32).tobytes())
    with pytest.raises(ValueError):
        io.read_flo(str(p))
    flow = np.zeros((2, 2, 2), np.float32)
    flow[1, 1, 0] = np.nan
    q = tmp_path / "nan.flo"
    io.write_flo(str(q), flow)
    with pytest.raises(ValueError):
        io.read_flo(str(q))


def test_read_flows_skips_missing(tmp_path):
    for i in (0, 2):
        io.write_flo(str(tmp_path / f"{i:05d}.flo"), np.full((3, 4, 2), i, np.float32))
    flows = io.read_flows(str(tmp_path), 3)
    assert [float(f[0, 0, 0]) for f in flows] == [0.0, 2.0]


def test_spix_csv_format(tmp_path):
    p = tmp_path / "00000.csv"
    io.save_spix_csv(str(p), np.array([[0, 1, 1], [2, 2, 10]], np.int32))
    assert p.read_text() == "0,1,1\n2,2,10\n"


def test_params_csv_format(tmp_path):
    n = 2
    params = {
        "mu_app": np.array([[0.5, 0.25, 1e-05], [1.0, 2.0, 3.0]], np.float32),
        "prior_mu_app": np.zeros((n, 3), np.float32),
        "mu_shape": np.array([[10.0, 20.5], [1234567.0, 0.1]]),
        "sigma_shape": np.array([[0.04, 0.0, 0.04], [1.0, -0.5, 2.0]]),
        "prior_sigma_shape": np.array([[0.04, 0.0, 0.04], [0.25, 0.0, 0.25]]),
        "counts": np.array([625, 3], np.int32),
        "prior_counts": np.array([100.0, 6.5], np.float32),
        "ids": np.array([0, 1], np.int32),
    }
    p = tmp_path / "00000_params.csv"
    io.save_params_csv(str(p), params)
    lines = p.read_text().splitlines()
    assert lines[0] == io.PARAMS_HEADER
    # std::ostream default formatting: 6 significant digits, no trailing zeros
    assert lines[1] == "0,0.5,0.25,1e-05,0,0,0,10,20.5,0.04,0,0.04,0.04,0,0.04,625,100"
    assert lines[2] == "1,1,2,3,0,0,0,1.23457e+06,0.1,1,-0.5,2,0.25,0,0.25,3,6.5"
