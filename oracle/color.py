"""CPU restatement of rgb2lab.cu:19-181 (TEST INFRASTRUCTURE, numpy): the
same double arithmetic with the power functions evaluated in float32, the
same channel conventions (B,G,R in; R,G,B out) and Lab scaling
(L/-100, a/100, b/100).  The float32 power is the host libm's, so results
match the device to within an ulp of powf, not bit-exactly."""
import numpy as np


def _powf(x, e):
    with np.errstate(invalid="ignore"):  # negative bases: the other branch is taken
        return np.power(np.asarray(x, np.float64).astype(np.float32),
                        np.float32(e)).astype(np.float64)


def rescale_image(img_u8, scale=255.0):
    """rescale_image_kernel, rgb2lab.cu:165-181."""
    return (np.asarray(img_u8).astype(np.float32) / np.float32(scale)).astype(np.float32)


def rgb2lab(bgr):
    """rgb_to_lab, rgb2lab.cu:50-117."""
    bgr = np.asarray(bgr, np.float32)
    B, G, R = (bgr[..., c].astype(np.float64) for c in range(3))

    def lin(v):
        return np.where(v <= 0.04045, v / 12.92, _powf((v + 0.055) / 1.055, 2.4))

    r, g, b = lin(R), lin(G), lin(B)
    X = r * 0.4124564 + g * 0.3575761 + b * 0.1804375
    Y = r * 0.2126729 + g * 0.7151522 + b * 0.0721750
    Z = r * 0.0193339 + g * 0.1191920 + b * 0.9503041
    eps, kappa = 0.008856, 903.3

    def f(t):
        return np.where(t > eps, _powf(t, np.float32(1.0 / 3.0)), (kappa * t + 16.0) / 116.0)

    fx, fy, fz = f(X / 0.950456), f(Y / 1.0), f(Z / 1.088754)
    lval = (116.0 * fy - 16.0).astype(np.float32)
    aval = (500.0 * (fx - fy)).astype(np.float32)
    bval = (200.0 * (fy - fz)).astype(np.float32)
    out = np.stack([lval / np.float32(-100), aval / np.float32(100), bval / np.float32(100)], -1)
    return out.astype(np.float32)


def lab2rgb(lab):
    """lab_to_rgb, rgb2lab.cu:119-162."""
    lab = np.asarray(lab, np.float32)
    L = (lab[..., 0] * np.float32(-100)).astype(np.float64)
    La = (lab[..., 1] * np.float32(100)).astype(np.float64)
    Lb = (lab[..., 2] * np.float32(100)).astype(np.float64)
    fy = (L + 16) / 116
    fx = La / 500 + fy
    fz = fy - Lb / 200

    def finv(v):
        c = _powf(v, 3.0)
        return np.where(c > 0.008856, c, (v - 16.0 / 116.0) / 7.787)

    X, Y, Z = 0.950456 * finv(fx), 1.000 * finv(fy), 1.088754 * finv(fz)
    R = X * 3.2406 + Y * (-1.5372) + Z * (-0.4986)
    G = X * -0.9689 + Y * 1.8758 + Z * 0.0415
    B = X * 0.0557 + Y * (-0.2040) + Z * 1.0570

    def gam(v):
        with np.errstate(invalid="ignore"):
            return np.where(v > 0.0031308, 1.055 * _powf(v, np.float32(1.0 / 2.4)) - 0.055,
                            12.92 * v)

    def clamp(v):
        p = np.minimum(1.0, v).astype(np.float32)
        return np.maximum(0.0, p.astype(np.float64)).astype(np.float32)

    return np.stack([clamp(gam(R)), clamp(gam(G)), clamp(gam(B))], -1)
