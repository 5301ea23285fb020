"""Seeded synthetic images and video for the BASS/BIST workloads.

Images are Voronoi mosaics: uniformly drawn seed points, each region a
constant colour drawn uniformly in CIE-Lab (L in [20, 90], a, b in
[-60, 60]), converted to sRGB (D65, gamut-clipped to [0, 1]) and then
perturbed by i.i.d. Gaussian noise.  Video moves every region's seed with a
per-region affine motion (translation |v| <= vmax px/frame plus a small
linear term); the forward flow of frame t is the motion of the region that
owns each pixel at frame t, plus Gaussian flow noise.  Regions may leave the
frame and new regions enter from the border (``churn``).

All arrays are float32, channels-last: images (H, W, 3), flows (H, W, 2)
with (u, v) = (dx, dy), as the reference's drop-in expects.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.spatial import cKDTree


def lab_to_rgb(lab: np.ndarray) -> np.ndarray:
    """CIE-Lab (D65) to sRGB in [0, 1] (clipped)."""
    L, a, b = lab[..., 0], lab[..., 1], lab[..., 2]
    fy = (L + 16.0) / 116.0
    fx = fy + a / 500.0
    fz = fy - b / 200.0
    eps = 6.0 / 29.0

    def finv(t):
        return np.where(t > eps, t ** 3, 3 * eps * eps * (t - 4.0 / 29.0))

    X = 0.95047 * finv(fx)
    Y = 1.00000 * finv(fy)
    Z = 1.08883 * finv(fz)
    r = 3.2406 * X - 1.5372 * Y - 0.4986 * Z
    g = -0.9689 * X + 1.8758 * Y + 0.0415 * Z
    bb = 0.0557 * X - 0.2040 * Y + 1.0570 * Z
    rgb = np.stack([r, g, bb], -1)
    rgb = np.where(rgb > 0.0031308, 1.055 * np.clip(rgb, 0, None) ** (1 / 2.4) - 0.055,
                   12.92 * rgb)
    return np.clip(rgb, 0.0, 1.0)


def random_colors(rng: np.random.Generator, n: int) -> np.ndarray:
    lab = np.stack([rng.uniform(20, 90, n), rng.uniform(-60, 60, n),
                    rng.uniform(-60, 60, n)], -1)
    return lab_to_rgb(lab).astype(np.float32)


def _voronoi_labels(seeds: np.ndarray, H: int, W: int) -> np.ndarray:
    yy, xx = np.mgrid[0:H, 0:W]
    pts = np.stack([xx.ravel(), yy.ravel()], -1).astype(np.float64)
    _, lab = cKDTree(seeds).query(pts)
    return lab.reshape(H, W).astype(np.int64)


def voronoi_image(H: int, W: int, nregions: int = 40, noise: float = 0.03,
                  seed: int = 0) -> np.ndarray:
    """One (H, W, 3) float32 Voronoi mosaic image."""
    rng = np.random.Generator(np.random.PCG64(seed))
    seeds = np.stack([rng.uniform(0, W, nregions), rng.uniform(0, H, nregions)], -1)
    cols = random_colors(rng, nregions)
    lab = _voronoi_labels(seeds, H, W)
    img = cols[lab] + rng.normal(0.0, noise, (H, W, 3))
    return img.astype(np.float32)


def voronoi_video(T: int, H: int, W: int, nregions: int = 120, vmax: float = 4.0,
                  noise: float = 0.03, flow_noise: float = 0.5, churn: float = 0.0,
                  seed: int = 0):
    """Video (T, H, W, 3) and forward flows (T-1, H, W, 2), float32.

    flows[t] maps frame t to frame t+1 (pixel (x, y) of frame t moves by
    flows[t, y, x]).  ``churn`` is the per-frame probability that a region
    re-enters from a random border point with a new colour.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    seeds = np.stack([rng.uniform(0, W, nregions), rng.uniform(0, H, nregions)], -1)
    ang = rng.uniform(0, 2 * np.pi, nregions)
    spd = rng.uniform(0, vmax * 0.8, nregions)
    vel = np.stack([spd * np.cos(ang), spd * np.sin(ang)], -1)
    # small linear term; |A (x-c)| stays below 0.2*vmax over a region
    A = rng.normal(0, 0.002, (nregions, 2, 2))
    cols = random_colors(rng, nregions)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    vid = np.zeros((T, H, W, 3), np.float32)
    flows = np.zeros((max(T - 1, 0), H, W, 2), np.float32)
    for t in range(T):
        lab = _voronoi_labels(seeds, H, W)
        vid[t] = cols[lab] + rng.normal(0.0, noise, (H, W, 3))
        if t == T - 1:
            break
        dx = xx - seeds[lab, 0]
        dy = yy - seeds[lab, 1]
        u = vel[lab, 0] + A[lab, 0, 0] * dx + A[lab, 0, 1] * dy
        v = vel[lab, 1] + A[lab, 1, 0] * dx + A[lab, 1, 1] * dy
        mag = np.sqrt(u * u + v * v)
        scale = np.minimum(1.0, vmax / np.maximum(mag, 1e-9))
        u, v = u * scale, v * scale
        flows[t, ..., 0] = u + rng.normal(0, flow_noise, (H, W))
        flows[t, ..., 1] = v + rng.normal(0, flow_noise, (H, W))
        seeds = seeds + vel
        if churn > 0:
            out = (seeds[:, 0] < -W * 0.05) | (seeds[:, 0] > W * 1.05) | \
                  (seeds[:, 1] < -H * 0.05) | (seeds[:, 1] > H * 1.05)
            redo = out | (rng.uniform(0, 1, nregions) < churn)
            n = int(redo.sum())
            if n:
                # re-enter on a random border side: 0 left, 1 right, 2 top, 3 bottom
                side = rng.integers(0, 4, n)
                pos = rng.uniform(0, 1, n)
                sx = np.select([side == 0, side == 1], [0.0, W - 1.0], pos * W)
                sy = np.select([side == 2, side == 3], [0.0, H - 1.0], pos * H)
                seeds[redo] = np.stack([sx, sy], -1)
                cols[redo] = random_colors(rng, n)
                ang = rng.uniform(0, 2 * np.pi, n)
                spd = rng.uniform(0, vmax * 0.8, n)
                vel[redo] = np.stack([spd * np.cos(ang), spd * np.sin(ang)], -1)
        else:
            # bounce seeds at the border so regions stay in view
            for c, lim in ((0, W), (1, H)):
                lo = seeds[:, c] < 0
                hi = seeds[:, c] > lim - 1
                vel[lo | hi, c] *= -1
                seeds[:, c] = np.clip(seeds[:, c], 0, lim - 1)
    return vid, flows


def batch_images(n: int = 32, lo=(240, 320), hi=(720, 1280), nregions: int = 40,
                 seed: int = 0):
    """n independent images with H, W drawn uniformly in [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        H = int(rng.integers(lo[0], hi[0] + 1))
        W = int(rng.integers(lo[1], hi[1] + 1))
        out.append(voronoi_image(H, W, nregions, seed=seed * 1000 + i + 1))
    return out


def voronoi_video_torch(T: int, H: int, W: int, nregions: int = 120, vmax: float = 4.0,
                        noise: float = 0.03, flow_noise: float = 0.5, seed: int = 0,
                        device="cuda", churn: float = 0.0):
    """Same motion model as voronoi_video, generated on the GPU with torch
    (used by the benchmark and the long-video tests for large frames), with
    the same ``churn`` rule: regions leaving the frame, or chosen with
    probability ``churn`` per frame, re-enter from a random border point
    with a new colour and velocity.  Returns (vid, flows) on device."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    seeds = torch.stack([torch.rand(nregions, generator=g, dtype=torch.float64) * W,
                         torch.rand(nregions, generator=g, dtype=torch.float64) * H], -1)
    ang = torch.rand(nregions, generator=g, dtype=torch.float64) * 2 * math.pi
    spd = torch.rand(nregions, generator=g, dtype=torch.float64) * vmax * 0.8
    vel = torch.stack([spd * torch.cos(ang), spd * torch.sin(ang)], -1)
    A = torch.randn(nregions, 2, 2, generator=g, dtype=torch.float64) * 0.002
    rng = np.random.Generator(np.random.PCG64(seed))
    cols = torch.from_numpy(random_colors(rng, nregions)).to(device)
    seeds, vel, A = seeds.to(device), vel.to(device), A.to(device)
    yy, xx = torch.meshgrid(torch.arange(H, device=device, dtype=torch.float64),
                            torch.arange(W, device=device, dtype=torch.float64), indexing="ij")
    gd = torch.Generator(device=device).manual_seed(seed + 1)
    vid = torch.empty((T, H, W, 3), dtype=torch.float32, device=device)
    flows = torch.empty((max(T - 1, 0), H, W, 2), dtype=torch.float32, device=device)
    for t in range(T):
        lab = torch.empty((H, W), dtype=torch.long, device=device)
        step = max(1, (1 << 24) // (W * nregions))
        for r0 in range(0, H, step):
            r1 = min(H, r0 + step)
            dx = xx[r0:r1, :, None] - seeds[None, None, :, 0]
            dy = yy[r0:r1, :, None] - seeds[None, None, :, 1]
            lab[r0:r1] = torch.argmin(dx * dx + dy * dy, -1)
        vid[t] = cols[lab] + noise * torch.randn((H, W, 3), generator=gd, device=device)
        if t == T - 1:
            break
        dx = xx - seeds[lab, 0]
        dy = yy - seeds[lab, 1]
        u = vel[lab, 0] + A[lab, 0, 0] * dx + A[lab, 0, 1] * dy
        v = vel[lab, 1] + A[lab, 1, 0] * dx + A[lab, 1, 1] * dy
        mag = torch.sqrt(u * u + v * v)
        scale = torch.clamp(vmax / torch.clamp(mag, min=1e-9), max=1.0)
        flows[t, ..., 0] = (u * scale).float() + flow_noise * torch.randn(
            (H, W), generator=gd, device=device)
        flows[t, ..., 1] = (v * scale).float() + flow_noise * torch.randn(
            (H, W), generator=gd, device=device)
        seeds = seeds + vel
        if churn > 0:
            out = (seeds[:, 0] < -W * 0.05) | (seeds[:, 0] > W * 1.05) | \
                  (seeds[:, 1] < -H * 0.05) | (seeds[:, 1] > H * 1.05)
            redo = (out.cpu() | (torch.rand(nregions, generator=g, dtype=torch.float64) < churn))
            n = int(redo.sum())
            if n:
                side = torch.randint(0, 4, (n,), generator=g)
                pos = torch.rand(n, generator=g, dtype=torch.float64)
                sx = torch.where(side == 0, 0.0, torch.where(side == 1, W - 1.0, pos * W))
                sy = torch.where(side == 2, 0.0, torch.where(side == 3, H - 1.0, pos * H))
                idx = torch.nonzero(redo).flatten().to(device)
                seeds[idx] = torch.stack([sx, sy], -1).to(device)
                cols[idx] = torch.from_numpy(random_colors(rng, n)).to(device)
                ang = torch.rand(n, generator=g, dtype=torch.float64) * 2 * math.pi
                spd = torch.rand(n, generator=g, dtype=torch.float64) * vmax * 0.8
                vel[idx] = torch.stack([spd * torch.cos(ang), spd * torch.sin(ang)], -1).to(device)
        else:
            for c, lim in ((0, W), (1, H)):
                out = (seeds[:, c] < 0) | (seeds[:, c] > lim - 1)
                vel[out, c] *= -1
                seeds[:, c] = seeds[:, c].clamp(0, lim - 1)
    return vid, flows
