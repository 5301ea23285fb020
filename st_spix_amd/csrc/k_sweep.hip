// Fused boundary update: S full sweeps (border detection + the 4 checkerboard
// phases each, reference update_seg.cu:281-311) in one launch.
//
// Each workgroup owns a TW x TH output tile.  It stages the tile's labels plus
// a halo of R = 4*S pixels in LDS.  A phase only reads the 3x3 neighbourhood,
// so after each phase the region whose labels are still exact shrinks by one
// pixel; phase p of sweep s therefore updates every pixel within radius
// 4*(S-1-s) + (3-p) of the tile, and after the last phase the tile itself is
// exact.  Halo pixels are recomputed redundantly by neighbouring tiles (no
// inter-workgroup communication).  Labels are read from `seg_in` and written
// to `seg_out` (ping-pong), so no tile can observe a neighbour's output.
#include "kernels.h"

namespace stsp {

__device__ __forceinline__ void joint_rec(const float* imgC, int x, int y, const SweepRec& r,
                                          int k, float sigma2_app, float neigh_neq,
                                          float beta, float& best, int& best_label) {
  float res = -1000.0f;
  const float x0 = imgC[0] - r.mu[0];
  const float x1 = imgC[1] - r.mu[1];
  const float x2 = imgC[2] - r.mu[2];
  const int d0 = (int)((double)x - r.mx);
  const int d1 = (int)((double)y - r.my);
  res = res - (x0 * x0 + x1 * x1 + x2 * x2) / sigma2_app;
  res = res - (float)(d0 * d0) * r.isig[0] - (float)(d1 * d1) * r.isig[2] -
        (float)(2 * d0 * d1) * r.isig[1];
  res = res - r.logdet;
  res = res - beta * neigh_neq;
  res = res - (float)(100 * (r.count <= 10));
  if (res > best) {
    best = res;
    best_label = k;
  }
}

__device__ __forceinline__ int nbr_bits(const int nb[8], int label) {
  int num = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) num |= ((label == nb[i]) || (nb[i] == -1)) << i;
  return num;
}

template <int TW, int TH, int S, int NT>
__global__ void __launch_bounds__(NT)
    k_sweep_fused(const float* __restrict__ img, const int* __restrict__ seg_in,
                  int* __restrict__ seg_out, const SweepRec* __restrict__ rec, float sigma2_app,
                  float potts, int H, int W, int tiles_x, SimpleLut lut) {
  constexpr int R = 4 * S;
  constexpr int LW = TW + 2 * R, LH = TH + 2 * R;
  __shared__ int lab[LH * LW];
  __shared__ uint8_t bord[LH * LW];
  __shared__ uint32_t lutb[8];
  const int tid = threadIdx.x;
  if (tid < 8) lutb[tid] = lut.bits[tid];
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int x0 = tx * TW - R, y0 = ty * TH - R;  // global coords of LDS (0,0)
  for (int i = tid; i < LH * LW; i += NT) {
    int lx = i % LW, ly = i / LW;
    int gx = x0 + lx, gy = y0 + ly;
    lab[i] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? seg_in[gy * W + gx] : -1;
  }
  __syncthreads();
  for (int s = 0; s < S; s++) {
    // border flags at the start of this sweep (seg_utils.cu:164-195)
    {
      const int rb = 4 * (S - 1 - s) + 3;
      const int bx0 = R - rb, by0 = R - rb, bw = TW + 2 * rb, bh = TH + 2 * rb;
      for (int i = tid; i < bw * bh; i += NT) {
        int lx = bx0 + i % bw, ly = by0 + i / bw;
        int gx = x0 + lx, gy = y0 + ly;
        int li = ly * LW + lx;
        uint8_t b = 0;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
          if (gy < 1 || gx < 1 || gy >= H - 1 || gx >= W - 1) {
            b = 1;
          } else {
            int C = lab[li];
            b = (C != lab[li - LW]) || (C != lab[li + LW]) || (C != lab[li + 1]) ||
                (C != lab[li - 1]);
          }
        }
        bord[li] = b;
      }
      __syncthreads();
    }
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
      const int xm = p >> 1, ym = p & 1;  // (0,0),(0,1),(1,0),(1,1) as the reference
      const int rp = 4 * (S - 1 - s) + (3 - p);
      const int rx0 = R - rp, ry0 = R - rp, rw = TW + 2 * rp, rh = TH + 2 * rp;
      // first local coordinate with the phase's global parity
      const int px0 = rx0 + (((xm - (x0 + rx0)) % 2) + 2) % 2;
      const int py0 = ry0 + (((ym - (y0 + ry0)) % 2) + 2) % 2;
      const int nx = (rx0 + rw - px0 + 1) / 2, ny = (ry0 + rh - py0 + 1) / 2;
      for (int i = tid; i < nx * ny; i += NT) {
        const int lx = px0 + 2 * (i % nx), ly = py0 + 2 * (i / nx);
        const int gx = x0 + lx, gy = y0 + ly;
        if (gx < 0 || gx >= W || gy < 0 || gy >= H) continue;
        const int li = ly * LW + lx;
        if (!bord[li]) continue;
        const int C = lab[li];
        const int N = lab[li - LW], Sd = lab[li + LW], Wl = lab[li - 1], E = lab[li + 1];
        const int nb[8] = {lab[li - LW - 1], N, lab[li - LW + 1], Wl, E, lab[li + LW - 1], Sd,
                           lab[li + LW + 1]};
        const int bN = nbr_bits(nb, N), bW = nbr_bits(nb, Wl), bS = nbr_bits(nb, Sd),
                  bE = nbr_bits(nb, E);
        const bool ok = ((lutb[bN >> 5] >> (bN & 31)) & 1u) && ((lutb[bW >> 5] >> (bW & 31)) & 1u) &&
                        ((lutb[bS >> 5] >> (bS & 31)) & 1u) && ((lutb[bE >> 5] >> (bE & 31)) & 1u);
        if (!ok) continue;
        const int gidx = gy * W + gx;
        const float imgC[3] = {img[3 * gidx], img[3 * gidx + 1], img[3 * gidx + 2]};
        float best = -9999.0f;
        int out = C;
        if (N >= 0)
          joint_rec(imgC, gx, gy, rec[N], N, sigma2_app, (float)(8 - __popc(bN)), potts, best, out);
        if (Sd != N && Sd >= 0)
          joint_rec(imgC, gx, gy, rec[Sd], Sd, sigma2_app, (float)(8 - __popc(bS)), potts, best,
                    out);
        if (Wl != Sd && Wl != N && Wl >= 0)
          joint_rec(imgC, gx, gy, rec[Wl], Wl, sigma2_app, (float)(8 - __popc(bW)), potts, best,
                    out);
        if (E != Wl && E != Sd && E != N && E >= 0)
          joint_rec(imgC, gx, gy, rec[E], E, sigma2_app, (float)(8 - __popc(bE)), potts, best, out);
        if (rec[C].count <= 2 && out == C) {
          if (N >= 0) out = N;
          else if (E >= 0) out = E;
          else if (Sd >= 0) out = Sd;
          else if (Wl >= 0) out = Wl;
        }
        lab[li] = out;
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < TW * TH; i += NT) {
    int lx = R + i % TW, ly = R + i / TW;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx < W && gy < H) seg_out[gy * W + gx] = lab[ly * LW + lx];
  }
}

template <int TW, int TH, int S, int NT>
static void launch_t(const float* img, const int* in, int* out, const SweepRec* rec, float s2,
                     float potts, int H, int W, const SimpleLut& lut, hipStream_t s) {
  int tx = (W + TW - 1) / TW, ty = (H + TH - 1) / TH;
  k_sweep_fused<TW, TH, S, NT><<<tx * ty, NT, 0, s>>>(img, in, out, rec, s2, potts, H, W, tx, lut);
}

// Runs niters_seg sweeps; returns the buffer holding the result (seg_a or
// seg_b).  `shape` selects the tile/sweeps-per-launch variant.
int* launch_sweeps(const float* img, int* seg_a, int* seg_b, const SweepRec* rec,
                   int niters_seg, float sigma2_app, float potts, int H, int W,
                   const SimpleLut& lut, int shape, hipStream_t s) {
  int* in = seg_a;
  int* out = seg_b;
  int done = 0;
  while (done < niters_seg) {
    int left = niters_seg - done, used;
    if (shape == 2 && left >= 4) {
      launch_t<64, 32, 4, 256>(img, in, out, rec, sigma2_app, potts, H, W, lut, s);
      used = 4;
    } else if (shape >= 1 && left >= 2) {
      launch_t<32, 32, 2, 256>(img, in, out, rec, sigma2_app, potts, H, W, lut, s);
      used = 2;
    } else {
      launch_t<32, 16, 1, 64>(img, in, out, rec, sigma2_app, potts, H, W, lut, s);
      used = 1;
    }
    done += used;
    int* t = in;
    in = out;
    out = t;
  }
  return in;
}

}  // namespace stsp
