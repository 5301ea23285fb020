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

}  // namespace stsp
