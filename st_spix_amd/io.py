"""On-disk formats of the reference's command-line path: Middlebury ``.flo``
optical flow (file_io.cpp:27-95) and the per-frame label and parameter CSV
files written by demo.cu:411-451 (save_spix, save_params)."""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

FLO_TAG = 202021.25


def read_flo(path: str) -> np.ndarray:
    """readOpticalFlow (file_io.cpp:48-95): float32 tag 202021.25, int32
    width and height, then (u, v) float32 pairs in row-major order.  Returns
    (H, W, 2) float32.  The reference prints an error and returns a dummy
    on a bad tag and exits on NaN; both raise ValueError here."""
    with open(path, "rb") as f:
        tag = np.fromfile(f, np.float32, 1)
        if tag.size != 1 or tag[0] != np.float32(FLO_TAG):
            raise ValueError(f"{path}: not a .flo file (tag {tag[:1]})")
        wh = np.fromfile(f, np.int32, 2)
        if wh.size != 2 or wh[0] < 1 or wh[1] < 1:
            raise ValueError(f"{path}: bad size {wh}")
        w, h = int(wh[0]), int(wh[1])
        data = np.fromfile(f, np.float32, 2 * w * h)
    if data.size != 2 * w * h:
        raise ValueError(f"{path}: truncated flow data")
    if np.isnan(data).any():
        raise ValueError(f"{path}: NaN in flow")
    return data.reshape(h, w, 2)


def write_flo(path: str, flow: np.ndarray) -> None:
    """Writes (H, W, 2) flow in the format read_flo reads."""
    flow = np.ascontiguousarray(flow, np.float32)
    h, w, c = flow.shape
    if c != 2:
        raise ValueError("flow must be (H, W, 2)")
    with open(path, "wb") as f:
        np.array([FLO_TAG], np.float32).tofile(f)
        np.array([w, h], np.int32).tofile(f)
        flow.tofile(f)


def read_flows(directory: str, num: int) -> List[np.ndarray]:
    """read_flows (file_io.cpp:27-45): frames 00000.flo .. {num:05d}.flo of a
    directory, in order, skipping the ones that do not exist."""
    out = []
    for i in range(num + 1):
        path = os.path.join(directory, f"{i:05d}.flo")
        if os.path.exists(path):
            out.append(read_flo(path))
    return out


def save_spix_csv(path: str, spix) -> None:
    """save_spix: one line per image row, labels separated by commas."""
    a = np.asarray(spix.cpu() if hasattr(spix, "cpu") else spix).astype(np.int64)
    with open(path, "w") as f:
        for row in a:
            f.write(",".join(str(int(v)) for v in row))
            f.write("\n")


PARAMS_HEADER = ("label,mu_i.x,mu_i.y,mu_i.z,p_mu_i.x,p_mu_i.y,p_mu_i.z,mu_s.x,mu_s.y,"
                 "sigma_s.x,sigma_s.y,sigma_s.z,prior_icov.x,prior_icov.y,prior_icov.z,"
                 "count,prior_count")


def _g(v) -> str:
    # std::ostream default float formatting: 6 significant digits, %g style
    return "%g" % float(v)


def save_params_csv(path: str, params: Dict) -> None:
    """save_params: the header line, then one row per id >= 0 with the
    appearance mean, its prior, the shape mean, the (inverse) shape
    covariance, the stored prior shape covariance (headed prior_icov, as in
    the reference), the count and the prior count."""
    p = {k: np.asarray(v.cpu() if hasattr(v, "cpu") else v) for k, v in params.items()}
    with open(path, "w") as f:
        f.write(PARAMS_HEADER + "\n")
        for k in p["ids"].astype(np.int64):
            if k < 0:
                continue
            vals = [*p["mu_app"][k], *p["prior_mu_app"][k], *p["mu_shape"][k],
                    *p["sigma_shape"][k], *p["prior_sigma_shape"][k]]
            row = [str(int(k))] + [_g(v) for v in vals]
            row += [str(int(p["counts"][k])), _g(p["prior_counts"][k])]
            f.write(",".join(row) + "\n")
