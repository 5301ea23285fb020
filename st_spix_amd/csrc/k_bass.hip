// Per-pixel and per-label kernels of the single-image BASS loop:
// hexagonal initialisation, border detection, the checkerboard boundary
// update, deterministic sufficient statistics, the Normal-Inverse-Wishart
// parameter update, compaction and parameter export.
#include "kernels.h"

namespace stsp {

// ---------------------------------------------------------------------------
// hexagonal initialisation (reference init_seg.cu:20-81)
// The reference compares each pixel with every centre (O(npix*K)); the
// nearest centre always lies in the 5x5 block of hex cells around the pixel,
// so only those 25 are compared, with the same distance expression and the
// same tie rule (last index wins).
// ---------------------------------------------------------------------------
__global__ void k_hex_init(int* seg, int H, int W, int nx, int ny, double Hh, double w) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= H * W) return;
  int x = idx % W, y = idx / W;
  int r0 = (int)floor((double)y / (1.5 * Hh));
  double D2 = 1.79769313486231570815e+308;
  int best = -1;
  for (int r = r0 - 2; r <= r0 + 2; r++) {
    if (r < 0 || r >= ny) continue;
    double off = (r % 2 == 0) ? 0.5 * w : 0.0;
    int c0 = (int)floor(((double)x - off) / w);
    for (int c = c0 - 2; c <= c0 + 2; c++) {
      if (c < 0 || c >= nx) continue;
      int j = r * nx + c;
      double cx = (double)c * w;
      double cy = (double)r * 1.5 * Hh;
      if (r % 2 == 0) cx = cx + 0.5 * w;
      double dx = (x - cx), dy = (y - cy);
      double d2 = dx * dx + dy * dy;
      if (d2 < D2 || (d2 == D2 && j > best)) {
        D2 = d2;
        best = j;
      }
    }
  }
  seg[idx] = best;
}

void launch_hex_init(int* seg, int H, int W, int sp_size, hipStream_t s) {
  HexGeom g = hex_geometry(sp_size, W, H);
  k_hex_init<<<div_up((long long)H * W, kBlock), kBlock, 0, s>>>(seg, H, W, g.nx, g.ny, g.Hh,
                                                                 g.w);
}

HexGeom hex_geometry(int sp_size, int width, int height) {
  HexGeom g;
  g.Hh = sqrt((double)pow((double)sp_size, 2) / (1.5 * sqrt(3.0)));
  g.w = sqrt(3.0) * g.Hh;
  g.nx = (int)floor((double)width / g.w) + 2;
  g.ny = (int)floor((double)height / (1.5 * g.Hh)) + 2;
  return g;
}

// ---------------------------------------------------------------------------
// border detection (reference seg_utils.cu:164-195)
// ---------------------------------------------------------------------------
__global__ void k_border(const int* __restrict__ seg, uint8_t* __restrict__ border, int H,
                         int W) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= H * W) return;
  int x = idx % W, y = idx / W;
  uint8_t b = 0;
  if (y < 1 || x < 1 || y >= H - 1 || x >= W - 1) {
    b = 1;
  } else {
    int C = seg[idx];
    b = (C != seg[idx - W]) || (C != seg[idx + W]) || (C != seg[idx + 1]) || (C != seg[idx - 1]);
  }
  border[idx] = b;
}

void launch_border(const int* seg, uint8_t* border, int H, int W, hipStream_t s) {
  k_border<<<div_up((long long)H * W, kBlock), kBlock, 0, s>>>(seg, border, H, W);
}

// ---------------------------------------------------------------------------
// boundary update of one checkerboard phase (reference update_seg.cu:29-259)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void joint(const float* imgC, int x, int y,
                                      const SweepRec* __restrict__ rec, int k,
                                      float sigma2_app, float neigh_neq, float beta,
                                      float& best, int& best_label) {
  const SweepRec r = rec[k];
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

// simple-point test + count of differing neighbours (seg_utils.h:30-79)
__device__ __forceinline__ int nbr_check(const int nb[8], int label, uint32_t lut_bits[8],
                                         bool& valid) {
  int num = 0, diff = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int same = (label == nb[i]) || (nb[i] == -1);
    num |= same << i;
    diff += 1 - same;
  }
  valid = (lut_bits[num >> 5] >> (num & 31)) & 1u;
  return diff;
}

__global__ void k_sweep_phase(const float* __restrict__ img, int* __restrict__ seg,
                              const uint8_t* __restrict__ border,
                              const SweepRec* __restrict__ rec, float sigma2_app, float potts,
                              int H, int W, int xmod, int ymod, SimpleLut lut) {
  int hw = (W - xmod + 1) / 2;
  int hh = (H - ymod + 1) / 2;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hw * hh) return;
  int x = 2 * (i % hw) + xmod;
  int y = 2 * (i / hw) + ymod;
  int idx = y * W + x;
  if (!border[idx]) return;
  int C = seg[idx];
  int N = -1, S = -1, E = -1, Wl = -1, NE = -1, NW = -1, SE = -1, SW = -1;
  if (x > 0) Wl = seg[idx - 1];
  if (y > 0) N = seg[idx - W];
  if (x < W - 1) E = seg[idx + 1];
  if (y < H - 1) S = seg[idx + W];
  if (y > 0 && x < W - 1) NE = seg[idx - W + 1];
  if (y > 0 && x > 0) NW = seg[idx - W - 1];
  if (x < W - 1 && y < H - 1) SE = seg[idx + W + 1];
  if (x > 0 && y < H - 1) SW = seg[idx + W - 1];
  const int nb[8] = {NW, N, NE, Wl, E, SW, S, SE};
  bool vN, vW, vS, vE;
  int dN = nbr_check(nb, N, lut.bits, vN);
  int dW = nbr_check(nb, Wl, lut.bits, vW);
  int dS = nbr_check(nb, S, lut.bits, vS);
  int dE = nbr_check(nb, E, lut.bits, vE);
  if (!vN || !vW || !vS || !vE) return;
  const float imgC[3] = {img[3 * idx], img[3 * idx + 1], img[3 * idx + 2]};
  float best = -9999.0f;
  int lab = C;
  if (N >= 0) joint(imgC, x, y, rec, N, sigma2_app, (float)dN, potts, best, lab);
  if (S != N && S >= 0) joint(imgC, x, y, rec, S, sigma2_app, (float)dS, potts, best, lab);
  if (Wl != S && Wl != N && Wl >= 0)
    joint(imgC, x, y, rec, Wl, sigma2_app, (float)dW, potts, best, lab);
  if (E != Wl && E != S && E != N && E >= 0)
    joint(imgC, x, y, rec, E, sigma2_app, (float)dE, potts, best, lab);
  if (rec[C].count <= 2 && lab == C) {
    if (N >= 0) lab = N;
    else if (E >= 0) lab = E;
    else if (S >= 0) lab = S;
    else if (Wl >= 0) lab = Wl;
  }
  seg[idx] = lab;
}

void launch_update_seg(const float* img, int* seg, uint8_t* border, const SweepRec* rec,
                       int niters_seg, float sigma2_app, float potts, int H, int W,
                       const SimpleLut& lut, hipStream_t s) {
  for (int it = 0; it < niters_seg; it++) {
    launch_border(seg, border, H, W, s);
    for (int xm = 0; xm < 2; xm++)
      for (int ym = 0; ym < 2; ym++) {
        int hw = (W - xm + 1) / 2, hh = (H - ym + 1) / 2;
        k_sweep_phase<<<div_up((long long)hw * hh, kBlock), kBlock, 0, s>>>(
            img, seg, border, rec, sigma2_app, potts, H, W, xm, ym, lut);
      }
  }
  launch_border(seg, border, H, W, s);
}

// ---------------------------------------------------------------------------
// sufficient statistics (reference update_params.cu:54-130), rule R1:
// int64 fixed-point colour sums, exact position moments.
// ---------------------------------------------------------------------------
__global__ void k_stats_clear(LabelStats st, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  st.count[k] = 0;
  st.ninvalid[k] = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    st.S[3 * k + c] = 0;
    st.Q[3 * k + c] = 0;
  }
#pragma unroll
  for (int c = 0; c < 5; c++) st.X[5 * k + c] = 0;
  if (st.bbox) {
    st.bbox[4 * k] = 0x7fffffff;
    st.bbox[4 * k + 1] = -1;
    st.bbox[4 * k + 2] = 0x7fffffff;
    st.bbox[4 * k + 3] = -1;
  }
}

// One wave aggregates runs of equal labels along its 64 consecutive pixels
// before touching global memory: each lane adds its pixel into the run head
// through a segmented shuffle reduction, and only run heads issue atomics.
template <bool BBOX>
__global__ void k_stats_accum(const float* __restrict__ img, const int* __restrict__ seg,
                              const int* __restrict__ shifted, int npix, int W, LabelStats st,
                              int nlab) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int lane = threadIdx.x & 63;
  int k = -1;
  int64_t v[11];
#pragma unroll
  for (int j = 0; j < 11; j++) v[j] = 0;
  int cnt = 0, ninv = 0;
  int bx0 = 0x7fffffff, bx1 = -1, by0 = 0x7fffffff, by1 = -1;
  if (t < npix) {
    k = seg[t];
    if (k >= nlab) k = -1;
    if (k >= 0) {
      float r = img[3 * t], g = img[3 * t + 1], b = img[3 * t + 2];
      int64_t x = t % W, y = t / W;
      v[0] = fx_color(r);
      v[1] = fx_color(g);
      v[2] = fx_color(b);
      v[3] = fx_square(r);
      v[4] = fx_square(g);
      v[5] = fx_square(b);
      v[6] = x;
      v[7] = y;
      v[8] = x * x;
      v[9] = x * y;
      v[10] = y * y;
      cnt = 1;
      if (shifted) ninv = shifted[t] < 0;
      if (BBOX) {
        bx0 = bx1 = (int)x;
        by0 = by1 = (int)y;
      }
    }
  }
  // segmented inclusive suffix reduction towards the run head (lowest lane of
  // the run): lane l adds lane l+d if that lane has the same label and lies
  // in the same run.
  int kprev = __shfl_up(k, 1);
  bool head = (lane == 0) || (kprev != k);
  // run id = lane of run head
  unsigned long long heads = __ballot(head);
  // distance to the next head above this lane
  unsigned long long above = heads & ~((2ull << lane) - 1ull);
  int run_end = above ? (__ffsll((long long)above) - 1) : 64;  // first lane of next run
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int c2 = __shfl_down(cnt, d);
    int n2 = __shfl_down(ninv, d);
    int64_t w2[11];
#pragma unroll
    for (int j = 0; j < 11; j++) w2[j] = __shfl_down(v[j], d);
    int ax0 = 0, ax1 = 0, ay0 = 0, ay1 = 0;
    if (BBOX) {
      ax0 = __shfl_down(bx0, d);
      ax1 = __shfl_down(bx1, d);
      ay0 = __shfl_down(by0, d);
      ay1 = __shfl_down(by1, d);
    }
    if (lane + d < run_end) {
      cnt += c2;
      ninv += n2;
#pragma unroll
      for (int j = 0; j < 11; j++) v[j] += w2[j];
      if (BBOX) {
        bx0 = min(bx0, ax0);
        bx1 = max(bx1, ax1);
        by0 = min(by0, ay0);
        by1 = max(by1, ay1);
      }
    }
  }
  if (head && k >= 0) {
    atomicAdd(st.count + k, cnt);
    if (ninv) atomicAdd(st.ninvalid + k, ninv);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      atomic_add_i64(st.S + 3 * k + c, v[c]);
      atomic_add_i64(st.Q + 3 * k + c, v[3 + c]);
    }
#pragma unroll
    for (int c = 0; c < 5; c++) atomic_add_i64(st.X + 5 * k + c, v[6 + c]);
    if (BBOX) {
      atomicMin(st.bbox + 4 * k, bx0);
      atomicMax(st.bbox + 4 * k + 1, bx1);
      atomicMin(st.bbox + 4 * k + 2, by0);
      atomicMax(st.bbox + 4 * k + 3, by1);
    }
  }
}

void launch_stats(const float* img, const int* seg, const int* shifted, int npix, int W,
                  LabelStats st, int nlab, hipStream_t s, bool bbox) {
  if (!bbox) st.bbox = nullptr;
  k_stats_clear<<<div_up(nlab, kBlock), kBlock, 0, s>>>(st, nlab);
  if (bbox)
    k_stats_accum<true><<<div_up(npix, kBlock), kBlock, 0, s>>>(img, seg, shifted, npix, W, st, nlab);
  else
    k_stats_accum<false><<<div_up(npix, kBlock), kBlock, 0, s>>>(img, seg, shifted, npix, W, st,
                                                                 nlab);
}

// ---------------------------------------------------------------------------
// parameter updates (reference update_params.cu:141-438)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void write_rec(SweepRec* rec, const Params& P, int k) {
  SweepRec r;
  r.mu[0] = P.mu_app[3 * k];
  r.mu[1] = P.mu_app[3 * k + 1];
  r.mu[2] = P.mu_app[3 * k + 2];
  r.logdet = (float)P.logdet[k];
  r.isig[0] = (float)P.sigma_shape[3 * k];
  r.isig[1] = (float)P.sigma_shape[3 * k + 1];
  r.isig[2] = (float)P.sigma_shape[3 * k + 2];
  r.count = P.count[k];
  r.mx = P.mu_shape[2 * k];
  r.my = P.mu_shape[2 * k + 1];
  rec[k] = r;
}

// clear_fields + count copy + calc_simple_update (update_params.cu:54-88, 220-381)
__global__ void k_param_update(Params P, LabelStats st, SweepRec* rec, int nbuf, int sp_size,
                               int summary_mode) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  // clear_fields
  P.mu_app[3 * k] = P.mu_app[3 * k + 1] = P.mu_app[3 * k + 2] = 0.f;
  P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0.0;
  P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0.0;
  int count_int = st.count[k];
  P.count[k] = count_int;
  if (P.valid[k] != 0) {
    if (count_int <= 0) {
      P.valid[k] = 0;
    } else if (summary_mode) {
      // calc_summ_stats (update_params.cu:141-203)
      double count = count_int * 1.0;
      for (int c = 0; c < 3; c++) P.mu_app[3 * k + c] = (float)(fx_to_double(st.S[3 * k + c]) / count);
      double mx = (double)st.X[5 * k] / count, my = (double)st.X[5 * k + 1] / count;
      P.mu_shape[2 * k] = mx;
      P.mu_shape[2 * k + 1] = my;
      double sx = (double)st.X[5 * k + 2] / count - mx * mx;
      double sy = (double)st.X[5 * k + 3] / count - mx * my;
      double sz = (double)st.X[5 * k + 4] / count - my * my;
      double det = sx * sz - sy * sy;
      if (det <= 0) {
        sx = sx + 0.00001;
        sz = sz + 0.00001;
        det = sx * sz - sy * sy;
        if (det <= 0) det = 0.00001;
      }
      P.sigma_shape[3 * k] = sx;
      P.sigma_shape[3 * k + 1] = sy;
      P.sigma_shape[3 * k + 2] = sz;
      P.logdet[k] = stsp_det_log(det);
    } else {
      float pc = P.prior_count[k];
      double count = count_int * 1.0;
      double total_count = (double)count_int + (double)(pc * 50.0f);
      double pi[3];
      if (P.prop[k]) {
        // _add_sigma_smoothing(sample_sigma_shape, (int)pc, pc, sp_size)
        int cnt = (int)pc;
        double tc = 1. * (float)((float)cnt + pc * 50.0f);
        const double* in = P.sample_sigma_shape + 3 * k;
        pi[0] = ((double)(pc * (float)sp_size) + cnt * in[0]) / (tc + 3.0);
        pi[1] = (cnt * in[1]) / (tc + 3.0);
        pi[2] = ((double)(pc * (float)sp_size) + cnt * in[2]) / (tc + 3.0);
      } else {
        pi[0] = P.prior_icov[3 * k];
        pi[1] = P.prior_icov[3 * k + 1];
        pi[2] = P.prior_icov[3 * k + 2];
      }
      double pr_det = pi[0] * pi[2] - pi[1] * pi[1];
      pr_det = sqrt(pr_det);
      double ps0 = (double)pc / pr_det * pi[0];
      double ps1 = (double)pc / pr_det * pi[1];
      double ps2 = (double)pc / pr_det * pi[2];
      for (int c = 0; c < 3; c++) P.mu_app[3 * k + c] = (float)(fx_to_double(st.S[3 * k + c]) / count);
      double mx = (double)st.X[5 * k] / count, my = (double)st.X[5 * k + 1] / count;
      P.mu_shape[2 * k] = mx;
      P.mu_shape[2 * k + 1] = my;
      double sx = (double)st.X[5 * k + 2] - count * mx * mx;
      double sy = (double)st.X[5 * k + 3] - count * mx * my;
      double sz = (double)st.X[5 * k + 4] - count * my * my;
      sx = ((double)pc * ps0 + sx) / (total_count + 3.0);
      sy = ((double)pc * ps1 + sy) / (total_count + 3.0);
      sz = ((double)pc * ps2 + sz) / (total_count + 3.0);
      double det = sx * sz - sy * sy;
      if (det <= 0) {
        sx = sx + 0.00001;
        sz = sz + 0.00001;
        det = sx * sz - sy * sy;
        if (det <= 0) det = 0.00001;
      }
      P.sigma_shape[3 * k] = sz / det;
      P.sigma_shape[3 * k + 1] = -sy / det;
      P.sigma_shape[3 * k + 2] = sx / det;
      P.logdet[k] = stsp_det_log(det);
    }
  }
  if (rec) write_rec(rec, P, k);
}

void launch_param_update(Params P, LabelStats st, SweepRec* rec, int nbuf, int sp_size,
                         int summary_mode, hipStream_t s) {
  k_param_update<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, st, rec, nbuf, sp_size,
                                                         summary_mode);
}

// store_sample_sigma_shape_k (update_params.cu:396-438)
__global__ void k_store_sample(Params P, LabelStats st, int nbuf) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf || P.valid[k] == 0) return;
  double count = P.count[k] * 1.0;
  double mx = (double)st.X[5 * k] / count, my = (double)st.X[5 * k + 1] / count;
  P.mu_shape[2 * k] = mx;
  P.mu_shape[2 * k + 1] = my;
  if (count > 0) {
    P.sample_sigma_shape[3 * k] = (double)st.X[5 * k + 2] / count - mx * mx;
    P.sample_sigma_shape[3 * k + 1] = (double)st.X[5 * k + 3] / count - mx * my;
    P.sample_sigma_shape[3 * k + 2] = (double)st.X[5 * k + 4] / count - my * my;
  } else {
    for (int c = 0; c < 3; c++) P.sample_sigma_shape[3 * k + c] = P.prior_sigma_shape[3 * k + c];
  }
}

void launch_store_sample(Params P, LabelStats st, int nbuf, hipStream_t s) {
  k_store_sample<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, st, nbuf);
}

// ---------------------------------------------------------------------------
// first-frame parameter initialisation (init_sparams.cu:25-166, 195-286)
// ---------------------------------------------------------------------------
__global__ void k_bass_mark(Params P, int nspix, int nbuf, int sp_size) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  // zero everything the reference leaves to the allocator (rule R11)
  P.icount[k] = 0.f;
  P.sm_count[k] = 0;
  for (int c = 0; c < 3; c++) {
    P.prior_mu_app[3 * k + c] = 0.f;
    P.sample_sigma_shape[3 * k + c] = 0.0;
    P.prior_icov[3 * k + c] = 0.0;
  }
  P.prior_mu_shape[2 * k] = P.prior_mu_shape[2 * k + 1] = 0.0;
  P.logdet[k] = 0.0;
  P.prior_count[k] = 0.f;
  P.valid[k] = k < nspix ? 1 : 0;
  P.prop[k] = k < nspix ? 1 : 0;
  if (k >= nspix) {
    P.prior_sigma_shape[3 * k] = 1.0 / sp_size;
    P.prior_sigma_shape[3 * k + 1] = 0;
    P.prior_sigma_shape[3 * k + 2] = 1.0 / sp_size;
  } else {
    P.prior_sigma_shape[3 * k] = P.prior_sigma_shape[3 * k + 1] = P.prior_sigma_shape[3 * k + 2] = 0;
  }
}

__global__ void k_bass_init(Params P, int nspix, int nbuf, int sp_size, int npix) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  int count = (int)(npix / (1. * nspix));
  P.prior_sigma_shape[3 * k] = 1.0 / sp_size;
  P.prior_sigma_shape[3 * k + 1] = 0.;
  P.prior_sigma_shape[3 * k + 2] = 1.0 / sp_size;
  if (k < nspix) {
    P.valid[k] = 1;
    P.prior_count[k] = (float)count;
    for (int c = 0; c < 3; c++) {
      P.prior_mu_app[3 * k + c] = P.mu_app[3 * k + c];
      P.mu_app[3 * k + c] = 0.f;
    }
    P.prior_icov[3 * k] = 1;
    P.prior_icov[3 * k + 1] = 0;
    P.prior_icov[3 * k + 2] = 1;
    P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0;
    P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0;
    P.prop[k] = 0;
  } else {
    P.count[k] = 0;
    for (int c = 0; c < 3; c++) {
      P.mu_app[3 * k + c] = 0.f;
      P.prior_mu_app[3 * k + c] = 0.f;
    }
    P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0;
    P.prior_mu_shape[2 * k] = P.prior_mu_shape[2 * k + 1] = 0;
    P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0;
    P.logdet[k] = 0;
    P.prop[k] = 0;
    P.valid[k] = 0;
    P.prior_icov[3 * k] = 1;
    P.prior_icov[3 * k + 1] = 0;
    P.prior_icov[3 * k + 2] = 1;
  }
}

void launch_bass_params_init(Params P, int nspix, int nbuf, int sp_size, int npix,
                             LabelStats st, const float* img, const int* seg, int W,
                             hipStream_t s) {
  k_bass_mark<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, nspix, nbuf, sp_size);
  launch_stats(img, seg, nullptr, npix, W, st, nbuf, s);
  launch_param_update(P, st, nullptr, nbuf, sp_size, 1, s);
  k_bass_init<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, nspix, nbuf, sp_size, npix);
}

// ---------------------------------------------------------------------------
// scans and compaction (compact_spix.cu:41-122, 196-295)
// ---------------------------------------------------------------------------
// Single-workgroup exclusive scan of n ints (n up to a few million); writes
// the total to *total if non-null.
__global__ void __launch_bounds__(1024) k_scan1(const int* __restrict__ in, int* __restrict__ out,
                                                int n, int* total) {
  __shared__ int part[1024];
  int tid = threadIdx.x;
  int chunk = (n + 1023) / 1024;
  int b = tid * chunk, e = min(n, b + chunk);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[tid] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - s;
  for (int i = b; i < e; i++) {
    int v = in[i];
    out[i] = run;
    run += v;
  }
  if (tid == 1023 && total) *total = part[1023];
}

void launch_scan(const int* in, int* out, int n, int* total, hipStream_t s) {
  k_scan1<<<1, 1024, 0, s>>>(in, out, n, total);
}

__global__ void k_fill_i32(int* p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill(int* p, int v, int n, hipStream_t s) {
  if (n > 0) k_fill_i32<<<div_up(n, kBlock), kBlock, 0, s>>>(p, v, n);
}

__global__ void k_mark_alive(const int* __restrict__ seg, int npix, int* alive, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int k = seg[t];
  if (k >= 0 && k < n) alive[k] = 1;
}
void launch_mark_alive(const int* seg, int npix, int* alive, int n, hipStream_t s) {
  k_mark_alive<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, npix, alive, n);
}

// alive flags of ids >= prev become "new" flags; ids below prev are kept
__global__ void k_new_flags(const int* alive, int* flags, int n, int prev) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  flags[k] = (k >= prev && alive[k]) ? 1 : 0;
}

__global__ void k_remap(int* seg, int npix, const int* rank, int prev) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int k = seg[t];
  if (k >= prev) seg[t] = prev + rank[k];
}

// the parameter fields fill_new_params_from_old copies (compact_spix.cu:220-295)
struct CompactTmp {
  float mu_app[3], prior_mu_app[3];
  double mu_shape[2], sigma_shape[3], prior_mu_shape[2], prior_sigma_shape[3];
  double sample_sigma_shape[3], logdet;
  int count, valid;
  float prior_count;
};

__global__ void k_compact_gather(Params P, const int* flags, const int* rank, int n, int prev,
                                 CompactTmp* tmp) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !flags[k]) return;
  CompactTmp t;
  for (int c = 0; c < 3; c++) {
    t.mu_app[c] = P.mu_app[3 * k + c];
    t.prior_mu_app[c] = P.prior_mu_app[3 * k + c];
    t.sigma_shape[c] = P.sigma_shape[3 * k + c];
    t.prior_sigma_shape[c] = P.prior_sigma_shape[3 * k + c];
    t.sample_sigma_shape[c] = P.sample_sigma_shape[3 * k + c];
  }
  for (int c = 0; c < 2; c++) {
    t.mu_shape[c] = P.mu_shape[2 * k + c];
    t.prior_mu_shape[c] = P.prior_mu_shape[2 * k + c];
  }
  t.logdet = P.logdet[k];
  t.count = P.count[k];
  t.valid = P.valid[k];
  t.prior_count = P.prior_count[k];
  tmp[rank[k]] = t;
}

__global__ void k_compact_scatter(Params P, const CompactTmp* tmp, const int* num_new, int prev,
                                  int nbuf) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= *num_new || prev + d >= nbuf) return;
  int k = prev + d;
  const CompactTmp t = tmp[d];
  for (int c = 0; c < 3; c++) {
    P.mu_app[3 * k + c] = t.mu_app[c];
    P.prior_mu_app[3 * k + c] = t.prior_mu_app[c];
    P.sigma_shape[3 * k + c] = t.sigma_shape[c];
    P.prior_sigma_shape[3 * k + c] = t.prior_sigma_shape[c];
    P.sample_sigma_shape[3 * k + c] = t.sample_sigma_shape[c];
  }
  for (int c = 0; c < 2; c++) {
    P.mu_shape[2 * k + c] = t.mu_shape[c];
    P.prior_mu_shape[2 * k + c] = t.prior_mu_shape[c];
  }
  P.logdet[k] = t.logdet;
  P.count[k] = t.count;
  P.valid[k] = t.valid;
  P.prior_count[k] = t.prior_count;
}

void launch_compact(int* seg, int npix, Params* P, int n, int prev, int nbuf, int* alive,
                    int* flags, int* rank, int* num_new, void* tmp, hipStream_t s) {
  launch_fill(alive, 0, n, s);
  launch_mark_alive(seg, npix, alive, n, s);
  k_new_flags<<<div_up(n, kBlock), kBlock, 0, s>>>(alive, flags, n, prev);
  launch_scan(flags, rank, n, num_new, s);
  if (P) {
    k_compact_gather<<<div_up(n, kBlock), kBlock, 0, s>>>(*P, flags, rank, n, prev,
                                                          (CompactTmp*)tmp);
  }
  k_remap<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, npix, rank, prev);
  if (P) {
    k_compact_scatter<<<div_up(n, kBlock), kBlock, 0, s>>>(*P, (const CompactTmp*)tmp, num_new,
                                                           prev, nbuf);
  }
}

size_t compact_tmp_bytes(int n) { return sizeof(CompactTmp) * (size_t)(n > 0 ? n : 1); }

// ---------------------------------------------------------------------------
// parameter export (read_params, sparams_io.cu:189-286) and update_prior with
// invert=false (sparams_io.cu:583-639, rule R10)
// ---------------------------------------------------------------------------
__global__ void k_export(Params P, OutParams o, const int* nspix_ptr, const int* alive) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x;
  int nspix = *nspix_ptr;
  if (ix >= nspix || ix >= o.cap) return;
  bool upd = alive[ix] && P.count[ix] > 0;
  for (int c = 0; c < 3; c++) {
    o.mu_app[3 * ix + c] = P.mu_app[3 * ix + c];
    o.prior_mu_app[3 * ix + c] = upd ? P.mu_app[3 * ix + c] : P.prior_mu_app[3 * ix + c];
    o.sigma_shape[3 * ix + c] = P.sigma_shape[3 * ix + c];
    o.prior_sigma_shape[3 * ix + c] =
        upd ? P.sigma_shape[3 * ix + c] : P.prior_sigma_shape[3 * ix + c];
    o.sample_sigma_shape[3 * ix + c] = P.sample_sigma_shape[3 * ix + c];
  }
  for (int c = 0; c < 2; c++) {
    o.mu_shape[2 * ix + c] = P.mu_shape[2 * ix + c];
    o.prior_mu_shape[2 * ix + c] = upd ? P.mu_shape[2 * ix + c] : P.prior_mu_shape[2 * ix + c];
  }
  o.logdet_sigma_shape[ix] = (float)P.logdet[ix];
  o.counts[ix] = P.count[ix];
  o.invalid_counts[ix] = P.icount[ix];
  o.sm_counts[ix] = P.sm_count[ix];
  o.prior_counts[ix] = P.prior_count[ix];
  o.ids[ix] = ix;
}

void launch_export(Params P, OutParams o, const int* nspix, const int* alive, int nmax,
                   hipStream_t s) {
  k_export<<<div_up(nmax, kBlock), kBlock, 0, s>>>(P, o, nspix, alive);
}

}  // namespace stsp

namespace stsp {

// ---------------------------------------------------------------------------
// border overlay of get_marked_video (pyapi.cu:221-303) with
// find_border_pixels_end (seg_utils.cu:218-257)
// ---------------------------------------------------------------------------
__global__ void k_marked(const float* __restrict__ vid, const int* __restrict__ spix,
                         float c0, float c1, float c2, int H, int W, long long total,
                         float* __restrict__ marked) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int npix = H * W;
  int idx = (int)(t % npix);
  long long base = t - idx;
  int x = idx % W, y = idx / W;
  int C = spix[t];
  int N = y > 0 ? spix[base + idx - W] : -1;
  int Wl = x > 0 ? spix[base + idx - 1] : -1;
  int S = y < H - 1 ? spix[base + idx + W] : -1;
  int E = x < W - 1 ? spix[base + idx + 1] : -1;
  bool b = (N >= 0 && C > N) || (S >= 0 && C > S) || (E >= 0 && C > E) || (Wl >= 0 && C > Wl);
  if (b) {
    marked[3 * t] = c0;
    marked[3 * t + 1] = c1;
    marked[3 * t + 2] = c2;
  } else {
    for (int c = 0; c < 3; c++) marked[3 * t + c] = fminf(fmaxf(vid[3 * t + c], 0.0f), 1.0f);
  }
}

void launch_marked(const float* vid, const int* spix, const float* color, int T, int H, int W,
                   float* marked, hipStream_t s) {
  long long total = (long long)T * H * W;
  k_marked<<<(int)((total + kBlock - 1) / kBlock), kBlock, 0, s>>>(vid, spix, color[0], color[1],
                                                                   color[2], H, W, total, marked);
}

}  // namespace stsp
