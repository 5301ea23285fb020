"""Seeded inputs of the configuration workloads whose oracle results are
committed as golden digests (tests/golden/config_goldens.json, written by
tests/golden/make_config_goldens.py).  Shared by the generator and the GPU
tests so that both build bit-identical inputs."""
from __future__ import annotations

import hashlib

import numpy as np

from st_spix_amd import synth

FIELDS = ["mu_app", "prior_mu_app", "mu_shape", "sigma_shape", "logdet_sigma_shape",
          "prior_mu_shape", "prior_sigma_shape", "sample_sigma_shape", "counts",
          "sm_counts", "prior_counts", "ids", "invalid_counts"]

# 30-frame 480x854 video with 120 regions, 100 iterations per frame; the
# first 5 frames are checked against the oracle
CONFIG1 = dict(T=5, H=480, W=854, nregions=120, vmax=4.0, seed=100,
               cfg=dict(niters=100, sp_size=25, potts=0.5, prop_niters=100))
# 32 images, H, W uniform in 240..720 x 320..1280, config 0's generator and
# iteration count, superpixel size 20
CONFIG2 = dict(n=32, seed=0, cfg=dict(niters=500, sp_size=20, potts=10.0))
# two of the 16 serving streams (360x640 and 1080x1920), config 1's motion
# model and iterations; frames 0..3 are checked against the oracle
CONFIG3_SAMPLE = [dict(T=20, H=360, W=640, nregions=120, vmax=4.0, seed=3000, check=4),
                  dict(T=20, H=1080, W=1920, nregions=120, vmax=4.0, seed=3003, check=4)]
CONFIG3_CFG = dict(niters=100, sp_size=25, potts=0.5, prop_niters=100)
# the 1080x1920 serving stream over 10 frames with config 1's iterations: from
# its fifth frame on some superpixels grow past 128 x 128 pixel boxes, the
# workgroup split searches (all 10 frames are checked against the oracle)
CONFIG3_LONG = dict(T=10, H=1080, W=1920, nregions=120, vmax=4.0, seed=3003)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def params_digest(p) -> str:
    return digest(*[np.asarray(p[f]) for f in FIELDS])


def config1_inputs():
    c = CONFIG1
    return synth.voronoi_video(c["T"], c["H"], c["W"], nregions=c["nregions"], vmax=c["vmax"],
                               seed=c["seed"])


def config2_sizes(lo=(240, 320), hi=(720, 1280)):
    """The (H, W) draws of synth.batch_images."""
    rng = np.random.Generator(np.random.PCG64(CONFIG2["seed"]))
    out = []
    for _ in range(CONFIG2["n"]):
        H = int(rng.integers(lo[0], hi[0] + 1))
        W = int(rng.integers(lo[1], hi[1] + 1))
        out.append((H, W))
    return out


def config2_image(i):
    """Image i of synth.batch_images(32, seed=0), generated alone."""
    H, W = config2_sizes()[i]
    return synth.voronoi_image(H, W, 40, seed=CONFIG2["seed"] * 1000 + i + 1)


def config2_inputs():
    return [config2_image(i) for i in range(CONFIG2["n"])]


def config3_sample_inputs(i, T=None):
    c = CONFIG3_SAMPLE[i]
    return synth.voronoi_video(T or c["T"], c["H"], c["W"], nregions=c["nregions"],
                               vmax=c["vmax"], seed=c["seed"])


def config3_long_inputs():
    c = CONFIG3_LONG
    return synth.voronoi_video(c["T"], c["H"], c["W"], nregions=c["nregions"], vmax=c["vmax"],
                               seed=c["seed"])
