// Device-side definitions shared by the BASS/BIST kernels (gfx950).
//
// Layout in HBM (one frame workspace per stream):
//   per pixel   : labels i32, border u8, shifted/halves/d1/d2/list/parent i32
//   per label   : structure-of-arrays parameters (cap slots), int64 fixed-point
//                 sufficient statistics, and a 48-byte "sweep record" that packs
//                 exactly what the boundary update reads for one candidate label.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "detmath.h"

namespace stsp {

// Fixed-point scales of the deterministic reductions: colour and colour
// squares use 2^-32, flow uses 2^-24.  Integer addition is associative, so
// every accumulation order (tiles, atomics, batching) gives the same bits.
constexpr double kFxColor = 4294967296.0;
constexpr double kFxFlow = 16777216.0;

constexpr int kBlock = 256;

struct alignas(16) SweepRec {
  float mu[3];
  float logdet;  // logdet_sigma_shape narrowed to float
  float isig[3]; // inverse shape covariance narrowed to float (xx, xy, yy)
  int count;
  // ceil and floor of the mean position.  The mean is X/n for integer sums
  // (or 0), which is either an integer or at least 1/n away from one, so
  // trunc((double)x - mean) == x - ceil(mean) when x >= ceil(mean) and
  // x - floor(mean) otherwise, exactly, for integer x.
  int cx, fx, cy, fy;
};

// Per-label sufficient statistics.  S = colour sums, Q = colour square sums
// (2^-32 fixed point), positions exact.  The live labels' sums (Q included,
// which the merge reads) are kept exact between full passes by the sweeps'
// per-pixel moves and the split/merge combinations.
struct LabelStats {
  int* count;    // [cap]
  int64_t* S;    // [cap*3]
  int64_t* Q;    // [cap*3]
  int64_t* X;    // [cap*5]  X, Y, XX, XY, YY
  int* ninvalid; // [cap]
  int* bbox;     // [cap*4]  min x, max x, min y, max y (gathered on request)
};

// Working parameters (the reference's spix_params, structure of arrays).
struct Params {
  float* mu_app;           // 3
  float* prior_mu_app;     // 3
  double* mu_shape;        // 2
  double* sigma_shape;     // 3 (inverse covariance)
  double* prior_mu_shape;  // 2
  double* prior_sigma_shape;   // 3
  double* sample_sigma_shape;  // 3
  double* prior_icov;      // 3
  double* logdet;          // 1
  int* count;
  float* icount;
  int* sm_count;
  float* prior_count;
  int* valid;
  int* prop;
};

// Split/merge helper state, 2*cap+2 slots.
struct SmState {
  int* count;
  int64_t* S;      // 3 per slot
  int64_t* Q;      // 3 per slot
  int* ninvalid;
  int* count_f;
  double* b_n;     // 3
  double* b_n_f;   // 3
  double* marg;    // own marginal
  double* hasting;
  int* merge;
  int* remove;
  int* select;
  int* has_bn;
  int* pairs;      // 2 per slot
  int* winner;     // per target: max accepting k
  int* accept;
  int* rank;
  int* valid0;
  float* pc0;
};

struct NigConsts {
  double a0, b0, logb0, logpi, log2, log_alpha;
};

__device__ __forceinline__ int64_t fx_color(float v) {
  return __double2ll_rn((double)v * kFxColor);
}
__device__ __forceinline__ int64_t fx_square(float v) {
  double d = (double)v;
  return __double2ll_rn(d * d * kFxColor);
}
__device__ __forceinline__ double fx_to_double(int64_t s) {
  return (double)s * (1.0 / kFxColor);
}
__device__ __forceinline__ void atomic_add_i64(int64_t* p, int64_t v) {
  atomicAdd((unsigned long long*)p, (unsigned long long)v);
}

// Wave-wide reductions by DPP, result in lane 63.  Row shifts by 1, 2, 4, 8
// build prefix sums inside each 16-lane row; row_bcast:15 (rows 1, 3) and
// row_bcast:31 (rows 2, 3) carry the lower rows into the last lane.  Lanes a
// move cannot source keep the identity.  All 64 lanes must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_take(int v, int identity) {
  return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int64_t dpp_take64(int64_t v) {
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(uint64_t)v, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)((uint64_t)v >> 32), CTRL, ROW_MASK, 0xf, false);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
__device__ __forceinline__ int wave_sum_to63(int v) {
  v += dpp_take<0x111, 0xf>(v, 0);
  v += dpp_take<0x112, 0xf>(v, 0);
  v += dpp_take<0x114, 0xf>(v, 0);
  v += dpp_take<0x118, 0xf>(v, 0);
  v += dpp_take<0x142, 0xa>(v, 0);
  v += dpp_take<0x143, 0xc>(v, 0);
  return v;
}
__device__ __forceinline__ int64_t wave_sum64_to63(int64_t v) {
  v += dpp_take64<0x111, 0xf>(v);
  v += dpp_take64<0x112, 0xf>(v);
  v += dpp_take64<0x114, 0xf>(v);
  v += dpp_take64<0x118, 0xf>(v);
  v += dpp_take64<0x142, 0xa>(v);
  v += dpp_take64<0x143, 0xc>(v);
  return v;
}
__device__ __forceinline__ int wave_min_to63(int v) {
  const int id = 0x7fffffff;
  v = min(v, dpp_take<0x111, 0xf>(v, id));
  v = min(v, dpp_take<0x112, 0xf>(v, id));
  v = min(v, dpp_take<0x114, 0xf>(v, id));
  v = min(v, dpp_take<0x118, 0xf>(v, id));
  v = min(v, dpp_take<0x142, 0xa>(v, id));
  v = min(v, dpp_take<0x143, 0xc>(v, id));
  return v;
}
__device__ __forceinline__ int wave_max_to63(int v) {
  const int id = -0x7fffffff - 1;
  v = max(v, dpp_take<0x111, 0xf>(v, id));
  v = max(v, dpp_take<0x112, 0xf>(v, id));
  v = max(v, dpp_take<0x114, 0xf>(v, id));
  v = max(v, dpp_take<0x118, 0xf>(v, id));
  v = max(v, dpp_take<0x142, 0xa>(v, id));
  v = max(v, dpp_take<0x143, 0xc>(v, id));
  return v;
}

// Runs of equal keys inside one wave: returns true on the run head (lowest
// lane of the run) and sets run_end to the first lane after the run.
__device__ __forceinline__ bool wave_run(int key, int& run_end) {
  const int lane = threadIdx.x & 63;
  int kprev = __shfl_up(key, 1);
  bool head = (lane == 0) || (kprev != key);
  unsigned long long heads = __ballot(head);
  unsigned long long above = heads & ~((2ull << lane) - 1ull);
  run_end = above ? (__ffsll((long long)above) - 1) : 64;
  return head;
}

// Sum of v over the lanes of this lane's run, valid on the run head.
__device__ __forceinline__ int64_t run_sum64(int64_t v, int run_end) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int64_t w = __shfl_down(v, d);
    if (lane + d < run_end) v += w;
  }
  return v;
}

// lane i takes lane i-1's value, lane 0 takes `first` (DPP wave_shr:1)
__device__ __forceinline__ uint64_t lane_up(uint64_t v, uint64_t first) {
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)first, (int)(uint32_t)v, 0x138, 0xf,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(first >> 32), (int)(uint32_t)(v >> 32),
                                             0x138, 0xf, 0xf, false);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
// lane i takes lane i+1's value, lane 63 takes `last` (DPP wave_shl:1)
__device__ __forceinline__ uint64_t lane_down(uint64_t v, uint64_t last) {
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)last, (int)(uint32_t)v, 0x130, 0xf,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(last >> 32), (int)(uint32_t)(v >> 32),
                                             0x130, 0xf, 0xf, false);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}
// a 64-bit value of one lane, broadcast
__device__ __forceinline__ uint64_t lane_get(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace stsp
