// Device helpers shared by the batched refinement-loop kernels (k_batch.hip,
// k_sweep.hip, k_split.hip): the block -> frame map of a batch launch, the
// NIG log marginal and the per-kind block prologues.
#pragma once

#include "batch.h"

namespace stsp {

__device__ __forceinline__ int frame_of(const int* __restrict__ off, int nf, int b) {
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// NIG marginal (split_merge_prop.cu:2975-3019); lgamma terms cancel exactly.
__device__ __forceinline__ double nig_marg(const NigConsts& c, double n, const double* bn) {
  double a_n = c.a0 + n / 2.0;
  double num = a_n * c.logb0 + 0.5 * stsp_det_log(1.0 / n);
  double t = 0.5 * n * c.logpi + n * c.log2;
  double m = 0.0;
  for (int ch = 0; ch < 3; ch++) {
    double den = a_n * stsp_det_log(bn[ch]) + t;
    m = m + (num - den);
  }
  return m;
}

// per-pixel kernels: pixels t0 + j * kBlock for j < kPixPerThread
#define PIX_PROLOGUE                                         \
  const int f = frame_of(b.off.pix, b.nf, blockIdx.x);       \
  const FrameDesc& F = b.fr[f];                              \
  const int t0 = (blockIdx.x - b.off.pix[f]) * kPixBlock + threadIdx.x;

#define LAB_PROLOGUE                                         \
  const int f = frame_of(b.off.lab, b.nf, blockIdx.x);       \
  const FrameDesc& F = b.fr[f];                              \
  const int k = (blockIdx.x - b.off.lab[f]) * kBlock + threadIdx.x;

#define SM_PROLOGUE                                          \
  const int f = frame_of(b.off.sm, b.nf, blockIdx.x);        \
  const FrameDesc& F = b.fr[f];                              \
  const int k = (blockIdx.x - b.off.sm[f]) * kBlock + threadIdx.x;

// The parameters and sweep record of label k from its statistics: a function
// of that label's sums and prior fields alone, so a caller that changed only
// some labels may update only those.
__device__ __forceinline__ void param_update_one(const FrameDesc& F, int k, int summary_mode) {
  Params P = F.P;
  LabelStats st = F.st;
  const int sp_size = F.sp_size;
  P.mu_app[3 * k] = P.mu_app[3 * k + 1] = P.mu_app[3 * k + 2] = 0.f;
  P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0.0;
  P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0.0;
  const int count_int = st.count[k];
  P.count[k] = count_int;
  if (P.valid[k] != 0) {
    if (count_int <= 0) {
      P.valid[k] = 0;
    } else if (summary_mode) {
      double count = count_int * 1.0;
      for (int c = 0; c < 3; c++)
        P.mu_app[3 * k + c] = (float)(fx_to_double(st.S[3 * k + c]) / count);
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
        int cnt = (int)pc;  // _add_sigma_smoothing(sample_sigma_shape, (int)pc, pc, sp_size)
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
      for (int c = 0; c < 3; c++)
        P.mu_app[3 * k + c] = (float)(fx_to_double(st.S[3 * k + c]) / count);
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
  SweepRec r;
  r.mu[0] = P.mu_app[3 * k];
  r.mu[1] = P.mu_app[3 * k + 1];
  r.mu[2] = P.mu_app[3 * k + 2];
  r.logdet = (float)P.logdet[k];
  r.isig[0] = (float)P.sigma_shape[3 * k];
  r.isig[1] = (float)P.sigma_shape[3 * k + 1];
  r.isig[2] = (float)P.sigma_shape[3 * k + 2];
  r.count = P.count[k];
  const double mx = P.mu_shape[2 * k], my = P.mu_shape[2 * k + 1];
  r.cx = (int)ceil(mx);
  r.fx = (int)floor(mx);
  r.cy = (int)ceil(my);
  r.fy = (int)floor(my);
  F.rec[k] = r;
}


}  // namespace stsp
