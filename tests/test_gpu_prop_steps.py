"""The prop_cuda per-step ops on the GPU (stsp_shift_labels,
stsp_fill_missing through st_spix_amd.prop_cuda) against the oracle, bit
for bit: random label maps and flows, batches, collisions, holes of every
size including a 1080p frame with 10% holes, duplicate and out-of-image
entries in the missing list, and the round cap."""
import numpy as np
import pytest
import torch

from st_spix_amd import prop_cuda

pytestmark = pytest.mark.gpu


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 37, 53), (3, 120, 200), (1, 1080, 1920)])
def test_shift_labels_matches_oracle(orc, shape):
    B, H, W = shape
    rng = np.random.default_rng(H * W + B)
    n = max(1, (H * W) // 400)
    yy, xx = np.mgrid[0:H, 0:W]
    spix = np.stack([((yy // 20) * ((W + 19) // 20) + xx // 20) % n for _ in range(B)])
    spix = spix.astype(np.int32)
    spix[rng.random(spix.shape) < 0.01] = n + 3  # labels without a flow entry
    flow = rng.integers(-9, 10, (B, n, 2)).astype(np.int32)
    print(f"shift_labels {shape}", flush=True)
    sh, ct = prop_cuda.shift_labels(cuda(spix), cuda(flow))
    torch.cuda.synchronize()
    for b in range(B):
        rs, rc = orc.prop_shift_labels(spix[b], flow[b])
        np.testing.assert_array_equal(sh[b].cpu().numpy(), rs, f"batch {b} labels")
        np.testing.assert_array_equal(ct[b].cpu().numpy(), rc, f"batch {b} counts")


def hole_map(H, W, frac, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    nx = (W + 24) // 25
    lab = ((yy // 25) * nx + xx // 25).astype(np.int32)
    seg = lab.copy()
    # holes in blobs and scattered pixels
    seg[rng.random((H, W)) < frac] = -1
    for _ in range(max(1, H * W // 20000)):
        cy, cx, r = rng.integers(0, H), rng.integers(0, W), rng.integers(2, 30)
        seg[max(0, cy - r):cy + r, max(0, cx - r):cx + r] = -1
    n = int(lab.max()) + 1
    cen = np.stack([(np.arange(n) % nx) * 25 + 12.0, (np.arange(n) // nx) * 25 + 12.0], -1)
    cen = cen.astype(np.float32) + rng.normal(0, 3, (n, 2)).astype(np.float32)
    return seg, cen


@pytest.mark.parametrize("H,W,frac,brk", [(2, 3, 0.5, 2000), (40, 64, 0.3, 2000),
                                          (240, 320, 0.2, 2000), (240, 320, 0.2, 5),
                                          (1080, 1920, 0.1, 2000)])
def test_fill_missing_matches_oracle(orc, H, W, frac, brk):
    seg, cen = hole_map(H, W, frac, seed=H + W)
    miss = np.flatnonzero(seg.ravel() < 0).astype(np.int32)
    print(f"fill_missing {H}x{W} holes {miss.size} break {brk}", flush=True)
    out = prop_cuda.fill_missing(cuda(seg[None]), cuda(cen), cuda(miss[None]), brk)
    torch.cuda.synchronize()
    ref, rounds, neg = orc.prop_fill_missing(seg, cen, miss, brk)
    np.testing.assert_array_equal(out[0].cpu().numpy(), ref)


def test_fill_missing_list_edge_cases(orc):
    """Duplicate entries, entries outside the image, labels without a centre
    and an empty list."""
    seg, cen = hole_map(50, 70, 0.3, seed=9)
    miss = np.flatnonzero(seg.ravel() < 0).astype(np.int32)
    miss = np.concatenate([miss, miss[::3], [-5, 50 * 70, 10 ** 6]]).astype(np.int32)
    cen = cen[: len(cen) // 2]  # the upper half of the labels has no centre
    out = prop_cuda.fill_missing(cuda(seg[None]), cuda(cen), cuda(miss[None]), 2000)
    ref, _, _ = orc.prop_fill_missing(seg, cen, miss, 2000)
    np.testing.assert_array_equal(out[0].cpu().numpy(), ref)
    empty = np.zeros((1, 0), np.int32)
    out = prop_cuda.fill_missing(cuda(seg[None]), cuda(cen), cuda(empty), 2000)
    np.testing.assert_array_equal(out[0].cpu().numpy(), seg)


def test_prop_steps_reject_bad_inputs():
    s = torch.zeros((1, 4, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError):
        prop_cuda.shift_labels(s.float(), torch.zeros((1, 2, 2), dtype=torch.int32, device="cuda"))
    with pytest.raises(RuntimeError):
        prop_cuda.fill_missing(s.repeat(2, 1, 1), torch.zeros((3, 2), device="cuda"),
                               torch.zeros((1, 0), dtype=torch.int32, device="cuda"), 10)
