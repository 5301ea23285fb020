// Batched refinement-loop kernels: every launch covers all frames of a batch
// (see batch.h).  Semantics per frame are those of the reference kernels cited
// at each function.  Where the reference depends on atomic arrival order the
// rule here is order-independent:
//   - colour/position sums are exact int64 fixed-point additions;
//   - the merge winner of a target f is the largest accepting k (atomicMax);
//   - new split / relabel ids are max_spix+1+rank in ascending label order;
//   - remove steps read valid / prior_count / pairs from a snapshot;
//   - the relabel candidate is the lexicographic minimum of (distance, id).
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

#define PIX_PROLOGUE                                         \
  const int f = frame_of(b.off.pix, b.nf, blockIdx.x);       \
  const FrameDesc& F = b.fr[f];                              \
  const int t = (blockIdx.x - b.off.pix[f]) * kBlock + threadIdx.x;

#define LAB_PROLOGUE                                         \
  const int f = frame_of(b.off.lab, b.nf, blockIdx.x);       \
  const FrameDesc& F = b.fr[f];                              \
  const int k = (blockIdx.x - b.off.lab[f]) * kBlock + threadIdx.x;

#define SM_PROLOGUE                                          \
  const int f = frame_of(b.off.sm, b.nf, blockIdx.x);        \
  const FrameDesc& F = b.fr[f];                              \
  const int k = (blockIdx.x - b.off.sm[f]) * kBlock + threadIdx.x;

// ---------------------------------------------------------------------------
// statistics (update_params.cu:54-130), rule R1
// ---------------------------------------------------------------------------
__global__ void kb_stats_clear(Batch b, int bbox) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  LabelStats st = F.st;
  st.count[k] = 0;
  st.ninvalid[k] = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    st.S[3 * k + c] = 0;
    st.Q[3 * k + c] = 0;
  }
#pragma unroll
  for (int c = 0; c < 5; c++) st.X[5 * k + c] = 0;
  if (bbox) {
    st.bbox[4 * k] = 0x7fffffff;
    st.bbox[4 * k + 1] = -1;
    st.bbox[4 * k + 2] = 0x7fffffff;
    st.bbox[4 * k + 3] = -1;
  }
}

// which: 0 = live labels into st (limit nbuf); 1 = split halves into sst
// (limit nsm, ninvalid from the shifted map); 2 = live labels into sst.
//
// A workgroup covers a kStatTile x kStatTile tile: wave w its rows 8w..8w+7,
// lane l the four pixels at column 4*(l&7) of row l>>3.  For each distinct
// label of a wave (usually one to three) the lanes sum their pixels of that
// label, DPP folds the sums into lane 63, and lane 63 issues one set of
// atomics.  Position moments are summed relative to the tile origin in int32
// and widened exactly: sum x = n*x0 + sum dx, sum x^2 = n*x0^2 + 2*x0*sum dx
// + sum dx^2, sum xy = n*x0*y0 + x0*sum dy + y0*sum dx + sum dx*dy.
// SQ: also gather the colour squares (only the split and merge passes read
// them; the live-label statistics never do).
// SUMS = false: bounding boxes (and the halves copy) only, for the pass
// before a split, when the sums are already current
template <bool BBOX, bool SQ, bool SUMS = true>
__global__ void __launch_bounds__(256) kb_stats_accum(Batch b, int cur, int which) {
  const int f = frame_of(b.off.st, b.nf, blockIdx.x);
  const FrameDesc& F = b.fr[f];
  const int W = F.W, H = F.H;
  const int tile = blockIdx.x - b.off.st[f];
  const int tiles_x = (W + kStatTile - 1) / kStatTile;
  const int tx0 = (tile % tiles_x) * kStatTile, ty0 = (tile / tiles_x) * kStatTile;
  const int lane = threadIdx.x & 63;
  const int lx = 4 * (lane & 7), ly = 8 * (threadIdx.x >> 6) + (lane >> 3);
  const int y = ty0 + ly;
  const int* labs = which == 1 ? F.halves : F.seg[cur];
  LabelStats st = which == 0 ? F.st : F.sst;
  const int nlab = which == 0 ? F.nbuf : F.nsm;
  const int* shifted = (which == 1 && F.prop_mode) ? F.shifted : nullptr;
  int k[4];
  float col[4][3];
  int inv[4];
  unsigned pend = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int x = tx0 + lx + j;
    k[j] = -1;
    inv[j] = 0;
    col[j][0] = col[j][1] = col[j][2] = 0.f;
    if (x < W && y < H) {
      const int t = y * W + x;
      const int v = labs[t];
      // the pass before a split also starts the halves map (a copy of the
      // labels; the split search then writes the moving parts)
      if (BBOX) F.halves[t] = v;
      if (v >= 0 && v < nlab) {
        k[j] = v;
        pend |= 1u << j;
        if (SUMS) {
          const float* p = F.img + 3 * t;
          col[j][0] = p[0];
          col[j][1] = p[1];
          col[j][2] = p[2];
          if (shifted) inv[j] = shifted[t] < 0;
        }
      }
    }
  }
  while (true) {
    const uint64_t act = __ballot(pend != 0);
    if (act == 0) break;
    const int leader = __builtin_ctzll(act);
    const int first = (pend & 1u) ? k[0] : (pend & 2u) ? k[1] : (pend & 4u) ? k[2] : k[3];
    const int u = __builtin_amdgcn_readlane(first, leader);
    int cnt = 0, sdx = 0, sdy = 0, sxx = 0, sxy = 0, syy = 0, ninv = 0;
    int64_t S[3] = {0, 0, 0}, Q[3] = {0, 0, 0};
    int mnx = 0x7fffffff, mxx = -1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (((pend >> j) & 1u) && k[j] == u) {
        pend &= ~(1u << j);
        const int dx = lx + j;
        cnt += 1;
        if (SUMS) {
          sdx += dx;
          sdy += ly;
          sxx += dx * dx;
          sxy += dx * ly;
          syy += ly * ly;
          ninv += inv[j];
#pragma unroll
          for (int c = 0; c < 3; c++) {
            S[c] += fx_color(col[j][c]);
            if (SQ) Q[c] += fx_square(col[j][c]);
          }
        }
        if (BBOX) {
          mnx = min(mnx, tx0 + dx);
          mxx = max(mxx, tx0 + dx);
        }
      }
    }
    int mny = cnt ? y : 0x7fffffff, mxy = cnt ? y : -1;
    if (SUMS) {
      cnt = wave_sum_to63(cnt);
      sdx = wave_sum_to63(sdx);
      sdy = wave_sum_to63(sdy);
      sxx = wave_sum_to63(sxx);
      sxy = wave_sum_to63(sxy);
      syy = wave_sum_to63(syy);
      if (shifted) ninv = wave_sum_to63(ninv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        S[c] = wave_sum64_to63(S[c]);
        if (SQ) Q[c] = wave_sum64_to63(Q[c]);
      }
    }
    if (BBOX) {
      mnx = wave_min_to63(mnx);
      mxx = wave_max_to63(mxx);
      mny = wave_min_to63(mny);
      mxy = wave_max_to63(mxy);
    }
    if (lane == 63 && !SUMS && BBOX) {
      atomicMin(st.bbox + 4 * u, mnx);
      atomicMax(st.bbox + 4 * u + 1, mxx);
      atomicMin(st.bbox + 4 * u + 2, mny);
      atomicMax(st.bbox + 4 * u + 3, mxy);
    }
    if (lane == 63 && SUMS) {
      const int64_t n = cnt, X0 = tx0, Y0 = ty0;
      atomicAdd(st.count + u, cnt);
      if (ninv) atomicAdd(st.ninvalid + u, ninv);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        atomic_add_i64(st.S + 3 * u + c, S[c]);
        if (SQ) atomic_add_i64(st.Q + 3 * u + c, Q[c]);
      }
      int64_t* X = st.X + 5 * u;
      atomic_add_i64(X + 0, n * X0 + sdx);
      atomic_add_i64(X + 1, n * Y0 + sdy);
      atomic_add_i64(X + 2, n * X0 * X0 + 2 * X0 * sdx + sxx);
      atomic_add_i64(X + 3, n * X0 * Y0 + X0 * sdy + Y0 * sdx + sxy);
      atomic_add_i64(X + 4, n * Y0 * Y0 + 2 * Y0 * sdy + syy);
      if (BBOX) {
        atomicMin(st.bbox + 4 * u, mnx);
        atomicMax(st.bbox + 4 * u + 1, mxx);
        atomicMin(st.bbox + 4 * u + 2, mny);
        atomicMax(st.bbox + 4 * u + 3, mxy);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// parameter update (update_params.cu:54-88, 141-203, 220-381)
// ---------------------------------------------------------------------------
__global__ void kb_param_update(Batch b, int summary_mode) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
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

// store_sample_sigma_shape_k (update_params.cu:396-438)
__global__ void kb_store_sample(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  Params P = F.P;
  LabelStats st = F.st;
  if (P.valid[k] == 0) return;
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

// parameters from the statistics as they stand (kept current by the sweeps)
void b_param_only(const Batch& b, hipStream_t s) {
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

void b_update_params(const Batch& b, int cur, bool bbox, hipStream_t s) {
  kb_stats_clear<<<b.lab_blocks, kBlock, 0, s>>>(b, bbox ? 1 : 0);
  if (bbox)
    kb_stats_accum<true, true><<<b.st_blocks, 256, 0, s>>>(b, cur, 0);
  else
    kb_stats_accum<false, true><<<b.st_blocks, 256, 0, s>>>(b, cur, 0);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

__global__ void kb_bbox_clear(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  int* bb = F.st.bbox + 4 * k;
  bb[0] = 0x7fffffff;
  bb[1] = -1;
  bb[2] = 0x7fffffff;
  bb[3] = -1;
}

// Before a split: the label sums are current (the sweeps move them, split and
// merge combine them), so the pass gathers only the bounding boxes and
// starts the halves map; then the parameters are updated.
void b_prepare_split(const Batch& b, int cur, hipStream_t s) {
  kb_bbox_clear<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_stats_accum<true, false, false><<<b.st_blocks, 256, 0, s>>>(b, cur, 0);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

void b_param_summary(const Batch& b, int cur, hipStream_t s) {
  kb_stats_clear<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
  kb_stats_accum<false, false><<<b.st_blocks, 256, 0, s>>>(b, cur, 0);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 1);
}

void b_store_sample(const Batch& b, hipStream_t s) {
  kb_store_sample<<<b.lab_blocks, kBlock, 0, s>>>(b);
}

// ---------------------------------------------------------------------------
// fused boundary sweeps (update_seg.cu:29-311): S sweeps per launch in LDS
// tiles with a 4S halo (see k_sweep_fused's description in k_sweep.hip)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void joint_b(const float* imgC, int x, int y, const SweepRec& r, int k,
                                        float sigma2_app, float neigh_neq, float beta,
                                        float& best, int& best_label) {
  float res = -1000.0f;
  const float x0 = imgC[0] - r.mu[0];
  const float x1 = imgC[1] - r.mu[1];
  const float x2 = imgC[2] - r.mu[2];
  const int d0 = x >= r.cx ? x - r.cx : x - r.fx;  // == (int)((double)x - mean x)
  const int d1 = y >= r.cy ? y - r.cy : y - r.fy;
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

// Statistics moves of one tile, aggregated per label in LDS before the
// global atomics: a linear-probing table of kMoveSlots labels with the net
// change of count, colour sums, colour squares (int64) and tile-relative
// position moments (int32, widened exactly when flushed).
constexpr int kMoveSlots = 64;
struct MoveTable {
  long long col[kMoveSlots][6];  // S r,g,b then Q r,g,b
  int pos[kMoveSlots][5];        // sum dx, dy, dx^2, dx dy, dy^2
  int cnt[kMoveSlots];
  int key[kMoveSlots];           // label, or -1
};

// slot of `label` (inserted if absent), or -1 when the table is full
__device__ __forceinline__ int move_slot(MoveTable& T, int label) {
  int h = (int)(((unsigned)label * 2654435761u) >> 26);  // 6 bits
#pragma unroll 1
  for (int probe = 0; probe < kMoveSlots; probe++) {
    const int prev = atomicCAS(&T.key[h], -1, label);
    if (prev == -1 || prev == label) return h;
    h = (h + 1) & (kMoveSlots - 1);
  }
  return -1;
}

// adds sign * (this pixel's statistics) to label's entry, or straight to
// the global sums when the table is full
__device__ __forceinline__ void move_add(MoveTable& T, const LabelStats& st, int label, int sign,
                                         const int64_t col[6], int dx, int dy, int64_t gx,
                                         int64_t gy) {
  const int h = move_slot(T, label);
  if (h >= 0) {
    atomicAdd(&T.cnt[h], sign);
#pragma unroll
    for (int c = 0; c < 6; c++)
      atomicAdd((unsigned long long*)&T.col[h][c], (unsigned long long)(sign * col[c]));
    atomicAdd(&T.pos[h][0], sign * dx);
    atomicAdd(&T.pos[h][1], sign * dy);
    atomicAdd(&T.pos[h][2], sign * dx * dx);
    atomicAdd(&T.pos[h][3], sign * dx * dy);
    atomicAdd(&T.pos[h][4], sign * dy * dy);
    return;
  }
  atomicAdd(st.count + label, sign);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    atomic_add_i64(st.S + 3 * label + c, sign * col[c]);
    atomic_add_i64(st.Q + 3 * label + c, sign * col[3 + c]);
  }
  const int64_t v[5] = {gx, gy, gx * gx, gx * gy, gy * gy};
#pragma unroll
  for (int c = 0; c < 5; c++) atomic_add_i64(st.X + 5 * label + c, sign * v[c]);
}

__device__ __forceinline__ int nbits(const int nb[8], int label) {
  int num = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) num |= ((label == nb[i]) || (nb[i] == -1)) << i;
  return num;
}

// One workgroup updates a TW x TH tile for S sweeps from an LDS copy with a
// halo of R = 4S.  Sweep s recomputes the border flags over the tile plus
// 4(S-1-s)+3 cells, and phase p of it updates the cells with
// (x%2, y%2) = (p>>1, p&1) within 4(S-1-s)+3-p cells of the tile: every
// cell of the written tile then sees exactly the labels a full-frame sweep
// would give it.  Border cells of each phase are compacted into LDS lists
// while the flags are computed, so a phase is one dense pass (about one cell
// per thread) with all its gathers in flight at once.  Cells of one phase are
// never 8-neighbours, so the order inside a list does not matter.
template <int TW, int TH, int S, int NT, int SHAPE>
__global__ void __launch_bounds__(NT) kb_sweep(Batch b, int cur, SimpleLut lut) {
  constexpr int R = 4 * S;
  constexpr int LW = TW + 2 * R, LH = TH + 2 * R;
  constexpr int RB0 = 4 * S - 1;  // widest border ring (sweep 0)
  constexpr int LCAP = ((TW + 2 * RB0 + 1) / 2) * ((TH + 2 * RB0 + 1) / 2);
  constexpr int LISTB = 4 * LCAP * 2;
  constexpr int POOLB = LISTB > (int)sizeof(MoveTable) ? LISTB : (int)sizeof(MoveTable);
  __shared__ int lab[LH * LW];
  // the phase lists, and after the last phase the table of statistics moves
  __shared__ __align__(16) unsigned char pool[POOLB];
  unsigned short (*list)[LCAP] = reinterpret_cast<unsigned short (*)[LCAP]>(pool);
  __shared__ int cnt[4];
  __shared__ uint32_t lutb[8];
  __shared__ uint32_t touched[(TW * TH + 31) / 32];  // tile cells whose label moved
  const int f = frame_of(b.off.tile[SHAPE], b.nf, blockIdx.x);
  const FrameDesc& F = b.fr[f];
  const int H = F.H, W = F.W;
  const float* __restrict__ img = F.img;
  const int* __restrict__ seg_in = F.seg[cur];
  int* __restrict__ seg_out = F.seg[cur ^ 1];
  const SweepRec* __restrict__ rec = F.rec;
  const float sigma2_app = F.sigma2_app, potts = F.potts;
  const int tiles_x = (W + TW - 1) / TW;
  const int tile = blockIdx.x - b.off.tile[SHAPE][f];
  const int tid = threadIdx.x;
  if (tid < 8) lutb[tid] = lut.bits[tid];
  for (int i = tid; i < (TW * TH + 31) / 32; i += NT) touched[i] = 0u;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int x0 = tx * TW - R, y0 = ty * TH - R;
  {
    // all loads of the tile are issued before the first LDS store: each
    // reads a clamped in-frame address, cells outside the frame get -1
    constexpr int NL = (LH * LW + NT - 1) / NT;
    int v[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int i = tid + k * NT;
      const int lx = i % LW, ly = i / LW;
      const int gx = x0 + lx, gy = y0 + ly;
      const int cx = min(max(gx, 0), W - 1), cy = min(max(gy, 0), H - 1);
      v[k] = seg_in[cy * W + cx];
    }
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int i = tid + k * NT;
      const int lx = i % LW, ly = i / LW;
      const int gx = x0 + lx, gy = y0 + ly;
      if (i < LH * LW) lab[i] = (gx >= 0 && gx < W && gy >= 0 && gy < H) ? v[k] : -1;
    }
  }
#pragma unroll
  for (int s = 0; s < S; s++) {  // unrolled: region widths are constants
    if (tid < 4) cnt[tid] = 0;
    __syncthreads();
    {
      // The region starts one cell inside the LDS tile (rb < R), so the four
      // neighbour reads of every cell are in the tile; they are issued
      // together (no short-circuit), which is faster than testing the frame
      // edges first.
      const int rb = 4 * (S - 1 - s) + 3;
      static_assert(4 * (S - 1) + 3 < R, "the flag region keeps one cell of margin");
      const int bx0 = R - rb, by0 = R - rb, bw = TW + 2 * rb, bh = TH + 2 * rb;
      for (int i = tid; i < bw * bh; i += NT) {
        const int lx = bx0 + i % bw, ly = by0 + i / bw;
        const int gx = x0 + lx, gy = y0 + ly;
        const int li = ly * LW + lx;
        const int C = lab[li];
        const bool diff = (C != lab[li - LW]) | (C != lab[li + LW]) | (C != lab[li + 1]) |
                          (C != lab[li - 1]);
        const bool edge = gy < 1 || gx < 1 || gy >= H - 1 || gx >= W - 1;
        const bool inside = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const int p = ((gx & 1) << 1) | (gy & 1);
        const int rp = 4 * (S - 1 - s) + (3 - p);
        const bool inreg = lx >= R - rp && lx < R + TW + rp && ly >= R - rp && ly < R + TH + rp;
        if (inside && (edge || diff) && inreg) {
          const int pos = atomicAdd(&cnt[p], 1);
          list[p][pos] = (unsigned short)li;
        }
      }
      __syncthreads();
    }
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
      const int n = cnt[p];
      for (int i = tid; i < n; i += NT) {
        const int li = list[p][i];
        const int lx = li % LW, ly = li / LW;
        const int gx = x0 + lx, gy = y0 + ly;
        const int C = lab[li];
        const int N = lab[li - LW], Sd = lab[li + LW], Wl = lab[li - 1], E = lab[li + 1];
        const int nb[8] = {lab[li - LW - 1], N, lab[li - LW + 1], Wl, E, lab[li + LW - 1], Sd,
                           lab[li + LW + 1]};
        // The neighbour pattern of each distinct label among N, S, W, E
        // (including -1 outside the frame) must pass the simple-point table.
        // Candidates are those labels >= 0 in the order N, S, W, E, each
        // once; `order` packs their indices (2 bits each).  The image value,
        // the own label's count and the first two candidates' records are
        // gathered before the table checks so their latency overlaps them.
        const int q[4] = {N, Sd, Wl, E};
        bool dup[4];
        int order = 0, ncand = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          dup[j] = false;
#pragma unroll
          for (int t = 0; t < j; t++) dup[j] = dup[j] || q[t] == q[j];
          if (!dup[j] && q[j] >= 0) {
            order |= j << (2 * ncand);
            ncand++;
          }
        }
        const int j0 = order & 3, j1 = (order >> 2) & 3;
        const int L0 = ncand > 0 ? (j0 == 0 ? N : j0 == 1 ? Sd : j0 == 2 ? Wl : E) : C;
        const int L1 = ncand > 1 ? (j1 == 1 ? Sd : j1 == 2 ? Wl : E) : L0;
        const SweepRec r0 = rec[L0], r1 = rec[L1];
        const int cntC = rec[C].count;
        const int gidx = gy * W + gx;
        const float imgC[3] = {img[3 * gidx], img[3 * gidx + 1], img[3 * gidx + 2]};
        // the four table words are read together (no short-circuit)
        int m[4];
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          m[j] = nbits(nb, q[j]);
          w[j] = lutb[m[j] >> 5];
        }
        uint32_t ok = 1u;
#pragma unroll
        for (int j = 0; j < 4; j++) ok &= (uint32_t)dup[j] | ((w[j] >> (m[j] & 31)) & 1u);
        if (!ok) continue;
        float best = -9999.0f;
        int out = C;
        if (ncand > 0) {
          const int mj = j0 == 0 ? m[0] : j0 == 1 ? m[1] : j0 == 2 ? m[2] : m[3];
          joint_b(imgC, gx, gy, r0, L0, sigma2_app, (float)(8 - __popc(mj)), potts, best, out);
        }
        if (ncand > 1) {
          const int mj = j1 == 1 ? m[1] : j1 == 2 ? m[2] : m[3];
          joint_b(imgC, gx, gy, r1, L1, sigma2_app, (float)(8 - __popc(mj)), potts, best, out);
        }
        for (int c = 2; c < ncand; c++) {
          const int j = (order >> (2 * c)) & 3;
          const int L = j == 2 ? Wl : E;
          const int mj = j == 2 ? m[2] : m[3];
          joint_b(imgC, gx, gy, rec[L], L, sigma2_app, (float)(8 - __popc(mj)), potts, best, out);
        }
        if (cntC <= 2 && out == C) {
          if (N >= 0) out = N;
          else if (E >= 0) out = E;
          else if (Sd >= 0) out = Sd;
          else if (Wl >= 0) out = Wl;
        }
        if (out != C) {
          lab[li] = out;
          const int cx = lx - R, cy = ly - R;
          if (cx >= 0 && cx < TW && cy >= 0 && cy < TH) {
            const int k = cy * TW + cx;
            atomicOr(&touched[k >> 5], 1u << (k & 31));
          }
        }
      }
      __syncthreads();
    }
  }
  // write the tile; a pixel that changed label moves its exact integer
  // statistics from the old label to the new one, so the label statistics
  // stay equal to a full recount (integer sums are order-free).  The moves
  // of the tile are summed per label in LDS (over the phase lists, no
  // longer needed) and each label's net change is added once.
  MoveTable& T = *reinterpret_cast<MoveTable*>(pool);
  for (int i = tid; i < kMoveSlots; i += NT) {
    T.key[i] = -1;
    T.cnt[i] = 0;
#pragma unroll
    for (int c = 0; c < 6; c++) T.col[i][c] = 0;
#pragma unroll
    for (int c = 0; c < 5; c++) T.pos[i][c] = 0;
  }
  __syncthreads();
  LabelStats st = F.st;
  const int X0 = x0 + R, Y0 = y0 + R;  // the tile's first pixel
  for (int i = tid; i < TW * TH; i += NT) {
    int lx = R + i % TW, ly = R + i / TW;
    int gx = x0 + lx, gy = y0 + ly;
    if (gx >= W || gy >= H) continue;
    const int g = gy * W + gx;
    const int nl = lab[ly * LW + lx];
    seg_out[g] = nl;
    if (!((touched[i >> 5] >> (i & 31)) & 1u)) continue;  // never moved: nl == seg_in[g]
    const int ol = seg_in[g];
    if (nl == ol) continue;
    // (the colour squares move too: the merge reads them)
    int64_t col[6];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      col[c] = fx_color(img[3 * g + c]);
      col[3 + c] = fx_square(img[3 * g + c]);
    }
    move_add(T, st, ol, -1, col, gx - X0, gy - Y0, gx, gy);
    move_add(T, st, nl, 1, col, gx - X0, gy - Y0, gx, gy);
  }
  __syncthreads();
  for (int i = tid; i < kMoveSlots; i += NT) {
    const int u = T.key[i];
    if (u < 0) continue;
    const int64_t n = T.cnt[i], sdx = T.pos[i][0], sdy = T.pos[i][1];
    if (n) atomicAdd(st.count + u, (int)n);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (T.col[i][c]) atomic_add_i64(st.S + 3 * u + c, T.col[i][c]);
      if (T.col[i][3 + c]) atomic_add_i64(st.Q + 3 * u + c, T.col[i][3 + c]);
    }
    const int64_t XX0 = X0, YY0 = Y0;
    const int64_t v[5] = {n * XX0 + sdx, n * YY0 + sdy,
                          n * XX0 * XX0 + 2 * XX0 * sdx + T.pos[i][2],
                          n * XX0 * YY0 + XX0 * sdy + YY0 * sdx + T.pos[i][3],
                          n * YY0 * YY0 + 2 * YY0 * sdy + T.pos[i][4]};
#pragma unroll
    for (int c = 0; c < 5; c++)
      if (v[c]) atomic_add_i64(st.X + 5 * u + c, v[c]);
  }
}

// runs niters_seg sweeps; returns the index of the buffer holding the result
int b_sweeps(const Batch& b, int cur, int niters_seg, int shape, const SimpleLut& lut,
             hipStream_t s) {
#define STSP_SWEEP(SH)                                                                   \
  kb_sweep<kSweepShape[SH].tw, kSweepShape[SH].th, kSweepShape[SH].s, kSweepShape[SH].nt, \
           SH><<<b.tile_blocks[SH], kSweepShape[SH].nt, 0, s>>>(b, cur, lut)
  int done = 0;
  while (done < niters_seg) {
    int sh = shape;
    if (kSweepShape[sh].s > niters_seg - done) sh = 0;
    switch (sh) {
      case 1: STSP_SWEEP(1); break;
      case 2: STSP_SWEEP(2); break;
      case 3: STSP_SWEEP(3); break;
      case 4: STSP_SWEEP(4); break;
      case 5: STSP_SWEEP(5); break;
      default: STSP_SWEEP(0); break;
    }
    done += kSweepShape[sh].s;
    cur ^= 1;
  }
  return cur;
#undef STSP_SWEEP
}

// Labels with at least one pixel after the last statistics pass, per frame,
// added to counter[slot] (the distinct labels of the map).
__global__ void kb_count_living(Batch b, int slot) {
  LAB_PROLOGUE
  const bool alive = k < F.nbuf && F.st.count[k] > 0;
  const int n = __popcll(__ballot(alive));
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(F.counter + slot, n);
}

void b_count_living(const Batch& b, int slot, hipStream_t s) {
  kb_count_living<<<b.lab_blocks, kBlock, 0, s>>>(b, slot);
}

// ---------------------------------------------------------------------------
// per-frame helpers
// ---------------------------------------------------------------------------
// one workgroup per frame: exclusive scan of n = nbuf entries; which selects
// (split accept -> sm.rank) or (relabel want_new -> rank); total -> F.total
__global__ void __launch_bounds__(1024) kb_scan(Batch b, int which) {
  const FrameDesc& F = b.fr[blockIdx.x];
  const int* in = which == 0 ? F.sm.accept : F.want_new;
  int* out = which == 0 ? F.sm.rank : F.rank;
  const int n = F.nbuf;
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int beg = tid * chunk, end = min(n, beg + chunk);
  int sum = 0;
  for (int i = beg; i < end; i++) sum += in[i];
  part[tid] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int i = beg; i < end; i++) {
    int v = in[i];
    out[i] = run;
    run += v;
  }
  if (tid == 1023) *F.total = part[1023];
}

__global__ void kb_reset_work(Batch b) {
  if (threadIdx.x < kWorkCounters && blockIdx.x == 0) b.work_n[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x < b.nf) b.fr[threadIdx.x].counter[0] = 0;
}

// ---------------------------------------------------------------------------
// split (split_merge_orig.cu:202-420, split_merge_prop.cu:184-395)
// ---------------------------------------------------------------------------
__global__ void kb_sm_clear(Batch b) {
  SM_PROLOGUE
  if (k >= F.nsm) return;
  LabelStats sst = F.sst;
  SmState sm = F.sm;
  sst.count[k] = 0;
  sst.ninvalid[k] = 0;
  for (int c = 0; c < 3; c++) {
    sst.S[3 * k + c] = 0;
    sst.Q[3 * k + c] = 0;
    sm.b_n[3 * k + c] = 0.0;
    sm.b_n_f[3 * k + c] = 0.0;
  }
  for (int c = 0; c < 5; c++) sst.X[5 * k + c] = 0;
  sm.count_f[k] = 0;
  sm.has_bn[k] = 0;
  sm.marg[k] = 0.0;
  sm.hasting[k] = -999999.0;
  sm.merge[k] = 0;
  sm.remove[k] = 0;
  sm.select[k] = 0;
  sm.accept[k] = 0;
  sm.winner[k] = -1;
  if (k < F.nbuf) {
    sm.pairs[2 * k] = -1;
    sm.pairs[2 * k + 1] = -1;
  }
}


// Labels that take part in the split (valid, non-empty, split not skipped),
// sorted by the size of their bounding box: boxes up to wave_box x wave_box
// (at most 128 x 128, the wave kernel's register rows) go to the wave list
// work[0..), boxes up to 448 x 384 to the block list work[work_cap..), larger
// ones to the big-block list work[2 work_cap..).

__global__ void kb_split_list(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  if (*F.max_spix > F.nbuf / 2 || F.P.valid[k] == 0 || F.P.count[k] <= 0) return;
  const int* bb = F.st.bbox + 4 * k;
  const int bw = bb[1] - bb[0] + 1, bh = bb[3] - bb[2] + 1;
  if (bw < 1 || bh < 1) return;
  // one atomic per list with a wave-uniform address, so the compiler
  // aggregates each wave's increments into one
  if (bw <= F.wave_box && bh <= F.wave_box) {
    const int pos = atomicAdd(b.work_n + 0, 1);
    if (pos < b.work_cap) b.work[pos] = (f << 24) | k;
  } else if (bw <= kBlockColsMax && bh <= kBlockRowsMax) {
    const int pos = atomicAdd(b.work_n + 2, 1);
    if (pos < b.work_cap) b.work[b.work_cap + pos] = (f << 24) | k;
  } else {
    const int pos = atomicAdd(b.work_n + 4, 1);
    if (pos < b.work_cap) b.work[2 * b.work_cap + pos] = (f << 24) | k;
  }
}

__device__ __forceinline__ int seed_of(const FrameDesc& F, const int* seg, int k, int dir) {
  int x, y;
  if (dir == 1 || dir == -1) {
    x = (int)(F.P.mu_shape[2 * k]) + dir;
    y = (int)(F.P.mu_shape[2 * k + 1]);
  } else {
    x = (int)(F.P.mu_shape[2 * k]);
    y = (int)(F.P.mu_shape[2 * k + 1]) + dir;
  }
  int ind = y * F.W + x;
  if (ind < 0 || ind > F.W * F.H - 1) return -1;
  return seg[ind] == k ? ind : -1;
}

// ---------------------------------------------------------------------------
// Two-seed split proposal (calc_split_candidate, split_merge.cu:757-800, and
// calc_seg_split 847-860) as a bit-parallel BFS.  The label's pixels inside
// its bounding box form a bitmask M (one 64-bit word per 64 columns of a
// row).  Level by level the visited sets grow as
//   N = (V<<1 | V>>1 | row above | row below) & M & ~V
// which reaches exactly the pixels at 4-connected distance L at level L.
// The reference marks a pixel for the new label when d1 > d2, with d = -1
// for pixels a seed never reaches.  That is: reached from seed 1, and not
// reached from seed 2 at a level no earlier than seed 1.  NK collects the
// pixels that seed 2 reaches while seed 1 already has them (d2 >= d1), so
// the new label takes V1 & ~NK.
// ---------------------------------------------------------------------------
template <int RC, int WD>
__device__ __forceinline__ void bits_expand(const uint64_t (&V)[RC][WD],
                                            const uint64_t (&M)[RC][WD], uint64_t (&N)[RC][WD],
                                            int lane) {
#pragma unroll
  for (int c = 0; c < RC; c++) {
#pragma unroll
    for (int w = 0; w < WD; w++) {
      const uint64_t v = V[c][w];
      uint64_t h = (v << 1) | (v >> 1);
      if (w > 0) h |= V[c][w - 1] >> 63;
      if (w + 1 < WD) h |= V[c][w + 1] << 63;
      const uint64_t last_prev = c > 0 ? lane_get(V[c > 0 ? c - 1 : 0][w], 63) : 0ull;
      const uint64_t first_next = c + 1 < RC ? lane_get(V[c + 1 < RC ? c + 1 : c][w], 0) : 0ull;
      const uint64_t up = lane_up(v, last_prev), dn = lane_down(v, first_next);
      N[c][w] = (h | up | dn) & M[c][w] & ~v;
    }
  }
}

// One wave: rows 64c + lane of a box of at most 64*RC rows and 64*WD columns
// live in registers.
template <int RC, int WD>
__device__ __forceinline__ void bfs_wave(const int* __restrict__ seg, int W, int k, int bx0,
                                         int by0, int bw, int bh, int s1, int s2, int k2,
                                         int* __restrict__ halves) {
  const int lane = threadIdx.x & 63;
  uint64_t M[RC][WD], V1[RC][WD], V2[RC][WD], NK[RC][WD];
#pragma unroll
  for (int c = 0; c < RC; c++) {
    const int nrows = min(64, bh - 64 * c);  // wave-uniform
#pragma unroll
    for (int w = 0; w < WD; w++) {
      M[c][w] = V1[c][w] = V2[c][w] = NK[c][w] = 0;
      const int col = 64 * w + lane;
      for (int r0 = 0; r0 < nrows; r0 += 16) {
        int v[16];  // 16 row loads in flight, then 16 ballots
#pragma unroll
        for (int j = 0; j < 16; j++)
          v[j] = (r0 + j < nrows && col < bw) ? seg[(by0 + 64 * c + r0 + j) * W + bx0 + col] : -1;
#pragma unroll
        for (int j = 0; j < 16; j++) {
          const uint64_t word = __ballot(v[j] == k);
          M[c][w] = lane == r0 + j ? word : M[c][w];
        }
      }
    }
  }
  // seeds (level 1)
  const int sy1 = s1 >= 0 ? s1 / W - by0 : -1, sx1 = s1 >= 0 ? s1 % W - bx0 : 0;
  const int sy2 = s2 >= 0 ? s2 / W - by0 : -1, sx2 = s2 >= 0 ? s2 % W - bx0 : 0;
#pragma unroll
  for (int c = 0; c < RC; c++) {
#pragma unroll
    for (int w = 0; w < WD; w++) {
      if (sy1 == 64 * c + lane && (sx1 >> 6) == w) V1[c][w] |= 1ull << (sx1 & 63);
      if (sy2 == 64 * c + lane && (sx2 >> 6) == w) V2[c][w] |= 1ull << (sx2 & 63);
    }
  }
  while (true) {
    uint64_t N1[RC][WD], N2[RC][WD];
    bits_expand<RC, WD>(V1, M, N1, lane);
    bits_expand<RC, WD>(V2, M, N2, lane);
    uint64_t grew = 0;
#pragma unroll
    for (int c = 0; c < RC; c++) {
#pragma unroll
      for (int w = 0; w < WD; w++) {
        V1[c][w] |= N1[c][w];
        NK[c][w] |= N2[c][w] & V1[c][w];
        V2[c][w] |= N2[c][w];
        grew |= N1[c][w] | N2[c][w];
      }
    }
    if (__ballot(grew != 0) == 0) break;  // wave-uniform
  }
#pragma unroll
  for (int c = 0; c < RC; c++) {
#pragma unroll
    for (int w = 0; w < WD; w++) {
      const uint64_t K = V1[c][w] & ~NK[c][w];
      const int col = 64 * w + lane;
      const int nrows = min(64, bh - 64 * c);
      for (int rr = 0; rr < nrows; rr++) {
        const uint64_t word = __shfl(K, rr);
        if (col < bw && ((word >> lane) & 1ull)) halves[(by0 + 64 * c + rr) * W + bx0 + col] = k2;
      }
    }
  }
}

// Persistent waves over the wave list, wave w taking items w, w + nwaves,
// ... (no shared cursor: thousands of waves would contend on one atomic).
// Loop control is wave-uniform and free of lane-conditional branches.
__global__ void __launch_bounds__(256) kb_bfs_wave(Batch b, int cur, int direction) {
  const int lane = threadIdx.x & 63;
  const int n_items = __builtin_amdgcn_readfirstlane(min(b.work_n[0], b.work_cap));
  const int gw = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int nwaves = gridDim.x * 4;
  for (int it = gw; it < n_items; it += nwaves) {
    const int wv = __builtin_amdgcn_readfirstlane(b.work[it]);
    const int f = wv >> 24, k = wv & 0xffffff;
    const FrameDesc& F = b.fr[f];
    const int* seg = F.seg[cur];
    const int* bb = F.st.bbox + 4 * k;
    const int bx0 = __builtin_amdgcn_readfirstlane(bb[0]);
    const int by0 = __builtin_amdgcn_readfirstlane(bb[2]);
    const int bw = __builtin_amdgcn_readfirstlane(bb[1]) - bx0 + 1;
    const int bh = __builtin_amdgcn_readfirstlane(bb[3]) - by0 + 1;
    const int s1 = __builtin_amdgcn_readfirstlane(seed_of(F, seg, k, direction));
    const int s2 = __builtin_amdgcn_readfirstlane(seed_of(F, seg, k, -direction));
    const int k2 = k + __builtin_amdgcn_readfirstlane(*F.max_spix) + 1;
    const int W = F.W;
    if (s1 < 0) continue;  // no pixel has d1 > d2: the label keeps all its pixels
    if (lane == 0) atomicAdd(F.search_count + kSearchWave, 1ull);
    if (bh <= 64) {
      if (bw <= 64) bfs_wave<1, 1>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, F.halves);
      else bfs_wave<1, 2>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, F.halves);
    } else {
      if (bw <= 64) bfs_wave<2, 1>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, F.halves);
      else bfs_wave<2, 2>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, F.halves);
    }
  }
}

// Boxes beyond the wave path: one workgroup of NW waves per label, in bands
// of rows.  Band b = NW*c + wave holds rows 56b - 4 .. 56b + 59 of the box in
// the 64 lanes of wave `wave` (chunk c), WD words per row in registers: lanes 4..59
// are its own rows, lanes 0..3 and 60..63 copies of the neighbouring bands'
// rows.  Four levels run on registers only (vertical neighbours by DPP wave
// shifts); the copies go stale from the outside in by one row per level, so
// the own rows stay exact.  Then the bands trade their outer own rows through
// LDS (double-buffered by round parity) and one barrier ends the round.  The
// search ends after a round in which no own row grew: the level before that
// round was already a fixed point.  The 256-thread kernel holds up to 8 bands
// (448 rows) of 6 words (384 columns); the 512-thread kernel for larger
// boxes 16 bands (896 rows) of 8 words (512 columns).
//
// A label whose box is larger is searched in a window of that size around
// seed 1 first.  Only pixels connected to seed 1 can move, so the window's
// result is the label's whenever the component of seed 1 ends inside the
// window: then every path between its pixels (and from seed 2, if seed 2 is
// in it) stays in the window, and a seed 2 outside the window lies in another
// component and reaches none of them.  If the component touches a window edge
// beyond which the label continues, nothing is written and the whole box is
// searched: with LDS bitsets up to kBlkWords words, else by per-pixel
// distances in global memory over the bounding box of each level's frontier.
constexpr int kHalo = 4, kOwn = 64 - 2 * kHalo;
constexpr int kBlockRows = 8 * kOwn, kBlockCols = 6 * 64;  // 256 threads, 2 chunks of 4 bands
constexpr int kBigRows = 16 * kOwn, kBigCols = 8 * 64;     // 512 threads, 2 chunks of 8 bands
static_assert(kSplitWinRows == kBigRows && kSplitWinCols == kBigCols,
              "the widest search window is the 512-thread band layout's");
static_assert(kBlockRowsMax == kBlockRows && kBlockColsMax == kBlockCols,
              "boxes routed to the 256-thread kernel fit its bands");
constexpr int kBlkWords = 2048;
constexpr int kBigWords = 18 * 512;  // whole-box words of the 512-thread kernel (147 KB of LDS)
enum { kEscLeft = 1, kEscRight = 2, kEscTop = 4, kEscBottom = 8 };

template <int NBANDS, int WDMAX, int NW>
struct BandEdge {
  uint64_t v[2][NBANDS][2][kHalo][2][WDMAX];  // [buf][band][top|bottom][row][seed][word]
  uint32_t g[2][NW];                          // [buf][wave]: own rows grew
  uint32_t esc;                               // seed 1's component leaves the window
};
typedef BandEdge<8, 6, 4> BlockEdge;
typedef BandEdge<16, 8, 8> BigEdge;

// Searches the window (bx0, by0, bw, bh) of label k.  esc_sides marks the
// window edges beyond which the label's box continues.  Returns true
// (block-uniform) when seed 1's component reaches such an edge with a pixel of
// label k just outside; the halves map is then left unchanged.
// With MLDS the label mask lives in LDS (mlds, [band][word][lane], 8 * 64 *
// NW * RB * WD bytes, lanes on consecutive words) instead of registers.
template <int NW, int RB, int WD, bool MLDS, class Edge>
__device__ __forceinline__ bool bfs_bands(const int* __restrict__ seg, int W, int k, int bx0,
                                          int by0, int bw, int bh, int s1, int s2, int k2,
                                          int* __restrict__ halves, Edge& E, int esc_sides,
                                          uint64_t* __restrict__ mlds) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint64_t M[MLDS ? 1 : RB][MLDS ? 1 : WD], V1[RB][WD], V2[RB][WD], NK[RB][WD];
  // the mask word of chunk c, word w for this lane
#define STSP_MASK(c, w) \
  (MLDS ? mlds[((NW * (c) + wv) * WD + (w)) * 64 + lane] : M[MLDS ? 0 : (c)][MLDS ? 0 : (w)])
  const bool own = lane >= kHalo && lane < 64 - kHalo;
  if (tid == 0) E.esc = 0;  // read after the round barriers below
#pragma unroll
  for (int c = 0; c < RB; c++) {
    const int row0 = kOwn * (NW * c + wv) - kHalo;  // row of lane 0 (wave-uniform)
#pragma unroll
    for (int w = 0; w < WD; w++) {
      uint64_t m = 0;
      V1[c][w] = V2[c][w] = NK[c][w] = 0;
      const int col = 64 * w + lane;
      if (row0 < bh) {
        for (int r0 = 0; r0 < 64; r0 += 16) {
          int v[16];  // 16 row loads in flight, then 16 ballots
#pragma unroll
          for (int j = 0; j < 16; j++) {
            const int y = row0 + r0 + j;
            v[j] = (y >= 0 && y < bh && col < bw) ? seg[(by0 + y) * W + bx0 + col] : -1;
          }
#pragma unroll
          for (int j = 0; j < 16; j++) {
            const uint64_t word = __ballot(v[j] == k);
            m = lane == r0 + j ? word : m;
          }
        }
      }
      if (MLDS)
        mlds[((NW * c + wv) * WD + w) * 64 + lane] = m;  // read back by this lane only
      else
        M[MLDS ? 0 : c][MLDS ? 0 : w] = m;
    }
  }
  {
    const int y1 = s1 / W - by0, x1 = s1 % W - bx0;
    const int y2 = s2 >= 0 ? s2 / W - by0 : -1000, x2 = s2 >= 0 ? s2 % W - bx0 : 0;
#pragma unroll
    for (int c = 0; c < RB; c++) {
      const int y = kOwn * (NW * c + wv) - kHalo + lane;
#pragma unroll
      for (int w = 0; w < WD; w++) {
        if (y1 == y && (x1 >> 6) == w) V1[c][w] |= 1ull << (x1 & 63);
        if (y2 == y && (x2 >> 6) == w) V2[c][w] |= 1ull << (x2 & 63);
      }
    }
  }
  // Cell (c, w) is word w of chunk c over the wave's 64 lanes.  A cell can
  // grow at a level only if it, or a horizontally adjacent cell of the same
  // chunk, changed at the level before (vertical neighbours are lanes of the
  // same cell, or halo rows replaced at a round's end, which marks the cell
  // too).  `act` (wave-uniform, bit c * WD + w) holds the cells that may
  // grow; skipped cells would compute N = 0, so results are unchanged.
  static_assert(RB * WD <= 32, "one activity bit per cell");
  constexpr uint32_t kRowBits = (1u << WD) - 1u;
  auto spread = [&](uint32_t ch) {
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < RB; c++) {
      const uint32_t r = (ch >> (c * WD)) & kRowBits;
      out |= ((r | (r << 1) | (r >> 1)) & kRowBits) << (c * WD);
    }
    return out;
  };
  uint32_t act = 0;
#pragma unroll
  for (int c = 0; c < RB; c++)
#pragma unroll
    for (int w = 0; w < WD; w++)
      if (__ballot((V1[c][w] | V2[c][w]) != 0) != 0) act |= 1u << (c * WD + w);
  act = spread(act);
  for (int round = 0;; round++) {
    uint32_t grew = 0;
#pragma unroll 1
    for (int j = 0; j < kHalo; j++) {
      uint32_t ch = 0;
#pragma unroll
      for (int c = 0; c < RB; c++) {
        uint64_t N1[WD], N2[WD];
#pragma unroll
        for (int w = 0; w < WD; w++) {
          N1[w] = N2[w] = 0;
          if (!((act >> (c * WD + w)) & 1u)) continue;  // wave-uniform
          uint64_t h1 = (V1[c][w] << 1) | (V1[c][w] >> 1), h2 = (V2[c][w] << 1) | (V2[c][w] >> 1);
          if (w > 0) {
            h1 |= V1[c][w - 1] >> 63;
            h2 |= V2[c][w - 1] >> 63;
          }
          if (w + 1 < WD) {
            h1 |= V1[c][w + 1] << 63;
            h2 |= V2[c][w + 1] << 63;
          }
          const uint64_t m = STSP_MASK(c, w);
          N1[w] = (h1 | lane_up(V1[c][w], 0ull) | lane_down(V1[c][w], 0ull)) & m & ~V1[c][w];
          N2[w] = (h2 | lane_up(V2[c][w], 0ull) | lane_down(V2[c][w], 0ull)) & m & ~V2[c][w];
        }
#pragma unroll
        for (int w = 0; w < WD; w++) {
          if (!((act >> (c * WD + w)) & 1u)) continue;  // N1 = N2 = 0 there
          V1[c][w] |= N1[w];
          NK[c][w] |= N2[w] & V1[c][w];
          V2[c][w] |= N2[w];
          const uint64_t any = N1[w] | N2[w];
          if (__ballot(any != 0) != 0) ch |= 1u << (c * WD + w);
          if (__ballot(own && any != 0) != 0) grew = 1;
        }
      }
      act = spread(ch);
    }
    // publish the outer own rows: lanes 4..7 (top) and 56..59 (bottom)
    const int wb = round & 1;
    const bool top = lane >= kHalo && lane < 2 * kHalo, bot = lane >= 64 - 2 * kHalo && own;
    if (top || bot) {
      const int side = bot, r = top ? lane - kHalo : lane - (64 - 2 * kHalo);
#pragma unroll
      for (int c = 0; c < RB; c++)
#pragma unroll
        for (int w = 0; w < WD; w++) {
          E.v[wb][NW * c + wv][side][r][0][w] = V1[c][w];
          E.v[wb][NW * c + wv][side][r][1][w] = V2[c][w];
        }
    }
    if (lane == 0) E.g[wb][wv] = grew;
    __syncthreads();
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) any |= E.g[wb][i];
    if (__builtin_amdgcn_readfirstlane(any) == 0) break;
    // lanes 0..3 take the band above's bottom rows, lanes 60..63 the band
    // below's top rows (none above band 0 or below the last band: zeros)
    const bool up = lane < kHalo, dn = lane >= 64 - kHalo;
    uint32_t ch = 0;
#pragma unroll
    for (int c = 0; c < RB; c++) {
      const int b = NW * c + wv;
      const int nb = up ? b - 1 : b + 1;
      const bool ok = (up || dn) && nb >= 0 && nb < NW * RB;
      const int r = up ? lane : lane - (64 - kHalo);
#pragma unroll
      for (int w = 0; w < WD; w++) {
        const uint64_t e1 = ok ? E.v[wb][ok ? nb : 0][up ? 1 : 0][ok ? r : 0][0][w] : 0ull;
        const uint64_t e2 = ok ? E.v[wb][ok ? nb : 0][up ? 1 : 0][ok ? r : 0][1][w] : 0ull;
        const uint64_t n1 = (up || dn) ? e1 : V1[c][w];
        const uint64_t n2 = (up || dn) ? e2 : V2[c][w];
        if (__ballot(n1 != V1[c][w] || n2 != V2[c][w]) != 0) ch |= 1u << (c * WD + w);
        V1[c][w] = n1;
        V2[c][w] = n2;
      }
    }
    // cells that changed at the last level or took new halo rows
    act |= spread(ch);
  }
  if (esc_sides) {
    // does seed 1's component touch an open window edge?  Loads outside the
    // window stay inside the label's box (esc_sides marks only such edges)
    bool esc = false;
#pragma unroll
    for (int c = 0; c < RB; c++) {
      const int y = kOwn * (NW * c + wv) - kHalo + lane;
      const bool mine = own && y < bh;
      if ((esc_sides & kEscLeft) && mine && (V1[c][0] & 1ull))
        esc = esc || seg[(by0 + y) * W + bx0 - 1] == k;
      if ((esc_sides & kEscRight) && mine) {
        const int wl = (bw - 1) >> 6;
        uint64_t word = V1[c][0];
#pragma unroll
        for (int w = 1; w < WD; w++) word = w == wl ? V1[c][w] : word;
        if ((word >> ((bw - 1) & 63)) & 1ull) esc = esc || seg[(by0 + y) * W + bx0 + bw] == k;
      }
      // top edge: window row 0 is own row kHalo of band 0; bottom edge: row
      // bh - 1 is in band (bh - 1) / kOwn
      const int band = NW * c + wv;
      const bool has_top = (esc_sides & kEscTop) && band == 0;
      const bool has_bot = (esc_sides & kEscBottom) && band == (bh - 1) / kOwn;
#pragma unroll
      for (int e = 0; e < 2; e++) {
        if (!(e == 0 ? has_top : has_bot)) continue;  // wave-uniform
        const int ln = e == 0 ? kHalo : kHalo + (bh - 1) % kOwn;
        const int gy = e == 0 ? by0 - 1 : by0 + bh;
#pragma unroll
        for (int w = 0; w < WD; w++) {
          const int col = 64 * w + lane;
          const int v = col < bw ? seg[gy * W + bx0 + col] : -1;
          const uint64_t outside = __ballot(v == k);
          const uint64_t inside = lane_get(V1[c][w], ln);
          esc = esc || (outside & inside) != 0;
        }
      }
    }
    if (__ballot(esc) != 0 && lane == 0) E.esc = 1;
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(E.esc)) return true;
  }
#pragma unroll
  for (int c = 0; c < RB; c++) {
    const int row0 = kOwn * (NW * c + wv) - kHalo;
#pragma unroll
    for (int w = 0; w < WD; w++) {
      const uint64_t K = V1[c][w] & ~NK[c][w];
      if (__ballot(own && K != 0) == 0) continue;  // wave-uniform: nothing moves here
      const int col = 64 * w + lane;
      for (int rr = kHalo; rr < 64 - kHalo; rr++) {
        const int y = row0 + rr;
        if (y >= bh) break;  // wave-uniform
        const uint64_t word = __shfl(K, rr);
        if (col < bw && ((word >> lane) & 1ull)) halves[(by0 + y) * W + bx0 + col] = k2;
      }
    }
  }
  return false;
#undef STSP_MASK
}

// Whole box, up to kBlkWords 64-column words: the same level recurrence over
// LDS bitsets, one word per thread and step, two barriers per level.
template <int NT>
__device__ __forceinline__ void bfs_lds(const int* __restrict__ seg, int W, int k, int bx0, int by0,
                                        int bw, int bh, int s1, int s2, int k2,
                                        int* __restrict__ halves, uint64_t* sM, uint64_t* sV1,
                                        uint64_t* sV2, uint64_t* sNK, int* changed) {
  const int tid = threadIdx.x;
  const int WD = (bw + 63) >> 6;
  const int nw = bh * WD;
  // mask: each wave ballots one 64-column word per step
  for (int p = tid; p < nw * 64; p += NT) {
    const int r = p / (WD * 64), col = p - r * (WD * 64);
    const bool bit = col < bw && seg[(by0 + r) * W + bx0 + col] == k;
    const uint64_t word = __ballot(bit);
    sM[p >> 6] = word;  // every lane writes the same word
    sV1[p >> 6] = 0;
    sV2[p >> 6] = 0;
    sNK[p >> 6] = 0;
  }
  if (tid < 2) changed[tid] = 0;
  __syncthreads();
  if (tid == 0) {
    if (s1 >= 0) {
      const int y = s1 / W - by0, x = s1 % W - bx0;
      sV1[y * WD + (x >> 6)] |= 1ull << (x & 63);
    }
    if (s2 >= 0) {
      const int y = s2 / W - by0, x = s2 % W - bx0;
      sV2[y * WD + (x >> 6)] |= 1ull << (x & 63);
    }
  }
  __syncthreads();
  for (int level = 0;; level++) {
    uint64_t n1[kBlkWords / NT], n2[kBlkWords / NT];
    int ch = 0;
#pragma unroll
    for (int j = 0; j < kBlkWords / NT; j++) {
      const int i = tid + NT * j;
      n1[j] = n2[j] = 0;
      if (i < nw) {
        const int r = i / WD, w = i - r * WD;
        const uint64_t m = sM[i], a = sV1[i], c = sV2[i];
        uint64_t e1 = (a << 1) | (a >> 1), e2 = (c << 1) | (c >> 1);
        if (w > 0) {
          e1 |= sV1[i - 1] >> 63;
          e2 |= sV2[i - 1] >> 63;
        }
        if (w + 1 < WD) {
          e1 |= sV1[i + 1] << 63;
          e2 |= sV2[i + 1] << 63;
        }
        if (r > 0) {
          e1 |= sV1[i - WD];
          e2 |= sV2[i - WD];
        }
        if (r + 1 < bh) {
          e1 |= sV1[i + WD];
          e2 |= sV2[i + WD];
        }
        n1[j] = e1 & m & ~a;
        n2[j] = e2 & m & ~c;
        ch |= (n1[j] | n2[j]) != 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBlkWords / NT; j++) {
      const int i = tid + NT * j;
      if (i < nw && (n1[j] | n2[j])) {
        const uint64_t v1 = sV1[i] | n1[j];
        sV1[i] = v1;
        sNK[i] |= n2[j] & v1;
        sV2[i] |= n2[j];
      }
    }
    // the flag of level L is read after the second barrier of level L and
    // cleared for level L + 1 before it; level L + 1's first barrier orders
    // that read before anyone clears it again
    if (ch) changed[level & 1] = 1;
    if (tid == 0) changed[(level + 1) & 1] = 0;
    __syncthreads();
    const int go = __builtin_amdgcn_readfirstlane(changed[level & 1]);
    if (!go) break;
  }
  for (int p = tid; p < nw * 64; p += NT) {
    const int r = p / (WD * 64), col = p - r * (WD * 64);
    const uint64_t K = sV1[p >> 6] & ~sNK[p >> 6];
    if (col < bw && ((K >> (p & 63)) & 1ull)) halves[(by0 + r) * W + bx0 + col] = k2;
  }
}

// Whole box of up to NT * J words for the 512-thread kernel: the label mask
// and the set NK stay in registers (thread t owns words t, t + NT, ...), the
// visited sets in LDS; the same level recurrence as bfs_lds.
template <int NT, int J>
__device__ __forceinline__ void bfs_lds_owned(const int* __restrict__ seg, int W, int k, int bx0,
                                              int by0, int bw, int bh, int s1, int s2, int k2,
                                              int* __restrict__ halves, uint64_t* sV1, uint64_t* sV2,
                                              int* changed) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int WD = (bw + 63) >> 6;
  const int nw = bh * WD;
  uint64_t M[J], NK[J];
#pragma unroll
  for (int j = 0; j < J; j++) {
    // lane l of a wave owns word i0 + l: the wave ballots those 64 words in turn
    const int i0 = 64 * (tid >> 6) + NT * j;  // wave-uniform
    uint64_t m = 0;
    if (i0 < nw) {
      for (int l0 = 0; l0 < 64; l0 += 16) {
        int v[16];  // 16 word loads in flight, then 16 ballots
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int i = i0 + l0 + q, r = i / WD, w = i - r * WD, col = 64 * w + lane;
          v[q] = (i < nw && col < bw) ? seg[(by0 + r) * W + bx0 + col] : -1;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const uint64_t word = __ballot(v[q] == k);
          m = lane == l0 + q ? word : m;
        }
      }
    }
    M[j] = m;
    NK[j] = 0;
    const int i = tid + NT * j;
    if (i < nw) {
      sV1[i] = 0;
      sV2[i] = 0;
    }
  }
  if (tid < 2) changed[tid] = 0;
  __syncthreads();
  if (tid == 0) {
    const int y = s1 / W - by0, x = s1 % W - bx0;
    sV1[y * WD + (x >> 6)] |= 1ull << (x & 63);
    if (s2 >= 0) {
      const int y2 = s2 / W - by0, x2 = s2 % W - bx0;
      sV2[y2 * WD + (x2 >> 6)] |= 1ull << (x2 & 63);
    }
  }
  __syncthreads();
  for (int level = 0;; level++) {
    uint64_t n1[J], n2[J];
    int ch = 0;
#pragma unroll
    for (int j = 0; j < J; j++) {
      const int i = tid + NT * j;
      n1[j] = n2[j] = 0;
      if (i < nw) {
        const int r = i / WD, w = i - r * WD;
        const uint64_t a = sV1[i], c = sV2[i];
        uint64_t e1 = (a << 1) | (a >> 1), e2 = (c << 1) | (c >> 1);
        if (w > 0) {
          e1 |= sV1[i - 1] >> 63;
          e2 |= sV2[i - 1] >> 63;
        }
        if (w + 1 < WD) {
          e1 |= sV1[i + 1] << 63;
          e2 |= sV2[i + 1] << 63;
        }
        if (r > 0) {
          e1 |= sV1[i - WD];
          e2 |= sV2[i - WD];
        }
        if (r + 1 < bh) {
          e1 |= sV1[i + WD];
          e2 |= sV2[i + WD];
        }
        n1[j] = e1 & M[j] & ~a;
        n2[j] = e2 & M[j] & ~c;
        ch |= (n1[j] | n2[j]) != 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; j++) {
      const int i = tid + NT * j;
      if (i < nw && (n1[j] | n2[j])) {
        const uint64_t v1 = sV1[i] | n1[j];
        sV1[i] = v1;
        NK[j] |= n2[j] & v1;
        sV2[i] |= n2[j];
      }
    }
    // flag protocol as in bfs_lds
    if (ch) changed[level & 1] = 1;
    if (tid == 0) changed[(level + 1) & 1] = 0;
    __syncthreads();
    const int go = __builtin_amdgcn_readfirstlane(changed[level & 1]);
    if (!go) break;
  }
  // the moving set K = V1 & ~NK goes to LDS (over V2) for coalesced stores
#pragma unroll
  for (int j = 0; j < J; j++) {
    const int i = tid + NT * j;
    if (i < nw) sV2[i] = sV1[i] & ~NK[j];
  }
  __syncthreads();
  for (int p = tid; p < nw * 64; p += NT) {
    const int r = p / (WD * 64), col = p - r * (WD * 64);
    if (col < bw && ((sV2[p >> 6] >> (p & 63)) & 1ull)) halves[(by0 + r) * W + bx0 + col] = k2;
  }
}

// Any box: level-synchronous BFS distances in global memory (d = -1 until
// reached, 1 at the seeds; a pixel of label k takes L + 1 when a 4-neighbour
// holds L).  A pixel can change at level L only next to a pixel set at level
// L - 1, so each level scans just the bounding box of the previous level's
// new pixels grown by one.  The box's frontier bounds rotate over three LDS
// slots: level L reads slot L % 3, collects into slot (L + 1) % 3 and clears
// slot (L + 2) % 3, whose last reader (level L - 1) finished before level L's
// start.  d1/d2 entries are written only at pixels of label k, so labels
// searched at the same time by other workgroups never share an entry.
template <int NT>
__device__ __forceinline__ void bfs_global(const int* __restrict__ seg, int W, int k, int bx0,
                                           int by0, int bw, int bh, int s1, int s2, int k2,
                                           int* __restrict__ halves, int* __restrict__ d1g,
                                           int* __restrict__ d2g, int (*fb)[4]) {
  const int tid = threadIdx.x, lane = tid & 63;
  const int nc = bw * bh;
  for (int i = tid; i < nc; i += NT) {
    const int g = (by0 + i / bw) * W + bx0 + i % bw;
    if (seg[g] != k) continue;
    d1g[g] = (g == s1) ? 1 : -1;
    d2g[g] = (g == s2) ? 1 : -1;
  }
  if (tid == 0) {
    const int x1 = s1 % W - bx0, y1 = s1 / W - by0;
    int mnx = x1, mxx = x1, mny = y1, mxy = y1;
    if (s2 >= 0) {
      const int x2 = s2 % W - bx0, y2 = s2 / W - by0;
      mnx = min(mnx, x2);
      mxx = max(mxx, x2);
      mny = min(mny, y2);
      mxy = max(mxy, y2);
    }
    fb[1][0] = mnx;
    fb[1][1] = mxx;
    fb[1][2] = mny;
    fb[1][3] = mxy;
    fb[2][0] = fb[2][2] = 0x7fffffff;
    fb[2][1] = fb[2][3] = -1;
  }
  __syncthreads();
  for (int level = 1;; level++) {
    const int* cur = fb[level % 3];
    const int fx0 = __builtin_amdgcn_readfirstlane(cur[0]);
    const int fx1 = __builtin_amdgcn_readfirstlane(cur[1]);
    const int fy0 = __builtin_amdgcn_readfirstlane(cur[2]);
    const int fy1 = __builtin_amdgcn_readfirstlane(cur[3]);
    if (fx1 < fx0) break;  // nothing was set at the previous level
    const int rx0 = max(fx0 - 1, 0), rx1 = min(fx1 + 1, bw - 1);
    const int ry0 = max(fy0 - 1, 0), ry1 = min(fy1 + 1, bh - 1);
    const int rw = rx1 - rx0 + 1, nr = rw * (ry1 - ry0 + 1);
    int mnx = 0x7fffffff, mxx = -1, mny = 0x7fffffff, mxy = -1;
    for (int i = tid; i < nr; i += NT) {
      const int lx = rx0 + i % rw, ly = ry0 + i / rw, g = (by0 + ly) * W + bx0 + lx;
      if (seg[g] != k) continue;
      const bool u1 = d1g[g] == -1, u2 = d2g[g] == -1;
      if (!u1 && !u2) continue;
      bool h1 = false, h2 = false;
      const int nb[4] = {ly > 0 ? g - W : -1, lx > 0 ? g - 1 : -1, ly < bh - 1 ? g + W : -1,
                         lx < bw - 1 ? g + 1 : -1};
      for (int j = 0; j < 4; j++) {
        const int q = nb[j];
        if (q < 0 || seg[q] != k) continue;
        h1 |= d1g[q] == level;
        h2 |= d2g[q] == level;
      }
      if (u1 && h1) d1g[g] = level + 1;
      if (u2 && h2) d2g[g] = level + 1;
      if ((u1 && h1) || (u2 && h2)) {
        mnx = min(mnx, lx);
        mxx = max(mxx, lx);
        mny = min(mny, ly);
        mxy = max(mxy, ly);
      }
    }
    mnx = wave_min_to63(mnx);
    mxx = wave_max_to63(mxx);
    mny = wave_min_to63(mny);
    mxy = wave_max_to63(mxy);
    int* nxt = fb[(level + 1) % 3];
    if (lane == 63 && mxx >= 0) {
      atomicMin(nxt + 0, mnx);
      atomicMax(nxt + 1, mxx);
      atomicMin(nxt + 2, mny);
      atomicMax(nxt + 3, mxy);
    }
    if (tid == 0) {
      int* clr = fb[(level + 2) % 3];
      clr[0] = clr[2] = 0x7fffffff;
      clr[1] = clr[3] = -1;
    }
    __syncthreads();
  }
  for (int i = tid; i < nc; i += NT) {
    const int g = (by0 + i / bw) * W + bx0 + i % bw;
    if (seg[g] == k && d1g[g] > d2g[g]) halves[g] = k2;
  }
}

// One workgroup of NW waves per label of work list LIST (1: boxes up to
// 448 x 384, NW = 4; 2: larger boxes, NW = 8).
template <int NW, int LIST, int MAXR, int MAXC, class Edge>
__device__ __forceinline__ void bfs_block_items(const Batch& b, int cur, int direction, Edge& edge,
                                                uint64_t* sM, uint64_t* sV1, uint64_t* sV2,
                                                uint64_t* sNK, int* changed, int (*fb)[4],
                                                int& item) {
  constexpr int NT = 64 * NW;
  const int tid = threadIdx.x;
  const int n_items = __builtin_amdgcn_readfirstlane(min(b.work_n[2 * LIST], b.work_cap));
  while (true) {
    __syncthreads();
    if (tid == 0) item = atomicAdd(b.work_n + 2 * LIST + 1, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);
    if (it >= n_items) break;
    const int wv = __builtin_amdgcn_readfirstlane(b.work[LIST * b.work_cap + it]);
    const int f = wv >> 24, k = wv & 0xffffff;
    const FrameDesc& F = b.fr[f];
    const int* seg = F.seg[cur];
    const int W = F.W;
    const int* bb = F.st.bbox + 4 * k;
    const int bx0 = __builtin_amdgcn_readfirstlane(bb[0]);
    const int by0 = __builtin_amdgcn_readfirstlane(bb[2]);
    const int bw = __builtin_amdgcn_readfirstlane(bb[1]) - bx0 + 1;
    const int bh = __builtin_amdgcn_readfirstlane(bb[3]) - by0 + 1;
    const int s1 = __builtin_amdgcn_readfirstlane(seed_of(F, seg, k, direction));
    const int s2 = __builtin_amdgcn_readfirstlane(seed_of(F, seg, k, -direction));
    const int k2 = k + __builtin_amdgcn_readfirstlane(*F.max_spix) + 1;
    int* halves = F.halves;
    if (s1 < 0) continue;  // no pixel has d1 > d2: the label keeps all its pixels
    // the band window: the whole box if it fits, else a window around seed 1
    const int win_h = min(MAXR, __builtin_amdgcn_readfirstlane(F.win_h));
    const int win_w = min(MAXC, __builtin_amdgcn_readfirstlane(F.win_w));
    int wx0 = bx0, wy0 = by0, ww = bw, wh = bh, sides = 0;
    if (bh > win_h) {
      wh = win_h;
      wy0 = min(max(s1 / W - win_h / 2, by0), by0 + bh - win_h);
      sides |= (wy0 > by0 ? kEscTop : 0) | (wy0 + wh < by0 + bh ? kEscBottom : 0);
    }
    if (bw > win_w) {
      ww = win_w;
      wx0 = min(max(s1 % W - win_w / 2, bx0), bx0 + bw - win_w);
      sides |= (wx0 > bx0 ? kEscLeft : 0) | (wx0 + ww < bx0 + bw ? kEscRight : 0);
    }
    int t2 = s2;
    if (t2 >= 0) {
      const int y = t2 / W, x = t2 % W;
      if (y < wy0 || y >= wy0 + wh || x < wx0 || x >= wx0 + ww) t2 = -1;
    }
    bool escaped;
#define STSP_BANDS(RB, WD)                                                                     \
  escaped = bfs_bands<NW, RB, WD, (NW > 4)>(seg, W, k, wx0, wy0, ww, wh, s1, t2, k2, halves, edge, \
                                            sides, sM)
    const int rb = wh <= NW * kOwn ? 1 : 2;
    if (NW == 4) {
      const int wd = ww <= 128 ? 2 : (ww <= 256 ? 4 : 6);
      if (rb == 1) {
        if (wd == 2) STSP_BANDS(1, 2);
        else if (wd == 4) STSP_BANDS(1, 4);
        else STSP_BANDS(1, 6);
      } else {
        if (wd == 2) STSP_BANDS(2, 2);
        else if (wd == 4) STSP_BANDS(2, 4);
        else STSP_BANDS(2, 6);
      }
    } else {
      const int wd = ww <= 256 ? 4 : 8;
      if (rb == 1) {
        if (wd == 4) STSP_BANDS(1, 4);
        else STSP_BANDS(1, 8);
      } else {
        if (wd == 4) STSP_BANDS(2, 4);
        else STSP_BANDS(2, 8);
      }
    }
#undef STSP_BANDS
    if (tid == 0)
      atomicAdd(F.search_count + (escaped ? kSearchWhole : sides ? kSearchWindow : kSearchBand), 1ull);
    if (!escaped) continue;
    const int nwords = bh * ((bw + 63) >> 6);
    if (nwords <= kBlkWords)
      bfs_lds<NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, sM, sV1, sV2, sNK, changed);
    else if (NW > 4 && nwords <= kBigWords)  // sM.. are one pool of 2 kBigWords here
      bfs_lds_owned<NT, kBigWords / NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, sM,
                                        sM + kBigWords, changed);
    else
      bfs_global<NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, F.d1, F.d2, fb);
  }
}

__global__ void __launch_bounds__(256) kb_bfs_block(Batch b, int cur, int direction) {
  __shared__ uint64_t sM[kBlkWords], sV1[kBlkWords], sV2[kBlkWords], sNK[kBlkWords];
  __shared__ BlockEdge edge;
  __shared__ int item, changed[2], fb[3][4];
  bfs_block_items<4, 1, kBlockRows, kBlockCols>(b, cur, direction, edge, sM, sV1, sV2, sNK, changed,
                                                fb, item);
}

// One LDS pool: the band mask (16 bands x 8 words x 64 lanes) and the band
// edges, or the bitsets of a whole-box search (up to kBigWords words), which
// runs only after the bands are done.
constexpr int kBigPool = 2 * kBigWords;
static_assert(16 * 8 * 64 <= 4 * kBlkWords && 4 * kBlkWords * 8 + sizeof(BigEdge) <= kBigPool * 8,
              "the band mask and edges fit the pool");
__global__ void __launch_bounds__(512) kb_bfs_big(Batch b, int cur, int direction) {
  __shared__ uint64_t pool[kBigPool];
  __shared__ int item, changed[2], fb[3][4];
  BigEdge& edge = *reinterpret_cast<BigEdge*>(pool + 4 * kBlkWords);
  bfs_block_items<8, 2, kBigRows, kBigCols>(b, cur, direction, edge, pool, pool + kBlkWords,
                                             pool + 2 * kBlkWords, pool + 3 * kBlkWords, changed,
                                             fb, item);
}

__global__ void kb_split_eval(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const int ms = *F.max_spix;
  const Params& P = F.P;
  if (ms > F.nbuf / 2 || P.valid[k] == 0) return;
  const int s = k + (ms + 1);
  if (s >= F.nbuf) return;
  const LabelStats& st = F.st;
  const LabelStats& sst = F.sst;
  const NigConsts& nc = F.nc;
  int count_f = P.count[k], count_k = sst.count[k], count_s = sst.count[s];
  if (count_f < 1 || count_k < 1 || count_s < 1) return;
  if (count_f != count_k + count_s) return;
  double bk[3], bs[3], bf[3];
  for (int c = 0; c < 3; c++) {
    double Qk = fx_to_double(sst.Q[3 * k + c]), Qs = fx_to_double(sst.Q[3 * s + c]);
    double Sk = fx_to_double(sst.S[3 * k + c]), Ss = fx_to_double(sst.S[3 * s + c]);
    double Sf = fx_to_double(st.S[3 * k + c]);
    bk[c] = nc.b0 + 0.5 * (Qk - (Sk * Sk) / (double)count_k);
    bs[c] = nc.b0 + 0.5 * (Qs - (Ss * Ss) / (double)count_s);
    bf[c] = nc.b0 + 0.5 * ((Qk + Qs) - (Sf * Sf) / (double)count_f);
  }
  double h = nig_marg(nc, (double)count_k, bk) + nig_marg(nc, (double)count_s, bs) -
             nig_marg(nc, (double)count_f, bf) + (double)F.split_alpha;
  if (F.shape_marg)  // extension, default off: the reference leaves this term unused
    h += stsp_split_shape_term(count_k, sst.X + 5 * k, count_s, sst.X + 5 * s,
                               P.prior_sigma_shape + 3 * k, (double)P.prior_count[k]);
  if (F.prop_mode) {
    int is_prop = P.prop[k];
    int nk = sst.ninvalid[k], ns = sst.ninvalid[s];
    int propogate_k = (nk <= ns);
    F.sm.select[k] = propogate_k || !is_prop;
    double iperc_k = nk / (1.0 * count_k);
    double iperc_s = nk / (1.0 * count_s);
    double iperc = propogate_k ? iperc_k : iperc_s;
    iperc = is_prop ? iperc : 1.0;
    h = h - 2.0 * (1.0 - iperc);
  }
  F.sm.hasting[k] = h;
  F.sm.accept[k] = h > -2.0 ? 1 : 0;
}

__global__ void kb_split_assign(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const int ms = *F.max_spix;
  if (ms > F.nbuf / 2 || !F.sm.accept[k]) return;
  Params P = F.P;
  const int s = ms + 1 + F.sm.rank[k];
  if (s >= F.nbuf) {  // no slot left for the new label: the split is dropped
    F.sm.accept[k] = 0;
    return;
  }
  F.sm.pairs[2 * k] = s;
  if (!F.prop_mode) {
    float pc = (float)fmax(P.prior_count[k] / 2.0, 8.0);
    P.prior_count[k] = pc;
    P.prior_count[s] = pc;
  } else {
    int nk = F.sst.ninvalid[k], ns = F.sst.ninvalid[k + ms + 1];
    int prop_k = (nk <= ns);
    float prior_count = P.prior_count[k];
    float half = (float)fmax(prior_count / 2.0, 36.0);
    float pcs;
    if (prop_k) {
      P.prior_count[k] = prior_count;
      pcs = (float)(prior_count / 2.);
    } else {
      P.prior_count[k] = half;
      pcs = half;
      P.prior_icov[3 * k] = pcs;
      P.prior_icov[3 * k + 1] = 0;
      P.prior_icov[3 * k + 2] = pcs;
    }
    P.prior_count[s] = pcs;
    P.prop[s] = 0;
    P.icount[k] = (float)(P.icount[k] / 2.0);
    P.prior_icov[3 * s] = pcs;
    P.prior_icov[3 * s + 1] = 0;
    P.prior_icov[3 * s + 2] = pcs;
    P.valid[s] = 1;
  }
}

__global__ void kb_split_apply(Batch b, int cur) {
  PIX_PROLOGUE
  if (t >= F.npix) return;
  const int ms = *F.max_spix;
  if (ms > F.nbuf / 2) return;
  int* seg = F.seg[cur];
  int k = seg[t];
  if (k < 0 || k >= F.nbuf || !F.sm.accept[k]) return;
  const int k2 = k + (ms + 1);
  const int h = F.halves[t];
  const bool take = F.prop_mode ? (F.sm.select[k] ? (h == k2) : (h == k)) : (h == k2);
  if (take) {
    int s = F.sm.pairs[2 * k];
    seg[t] = s;
    F.P.valid[s] = 1;
  }
}

// Label statistics after a split without a recount: the halves pass
// gathered both parts of every label (k and k + max_spix + 1) with exact
// integer sums, and the labels that did not split kept their pixels, so an
// accepted label takes the sums of the part that stays and its new label
// those of the part that moves (the assignment of kb_split_apply).  Must run
// before kb_add_max changes max_spix.  (The invalid counts are not kept for
// live labels.)
__global__ void kb_split_stats(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const int ms = *F.max_spix;
  if (ms > F.nbuf / 2 || !F.sm.accept[k]) return;
  const int s = F.sm.pairs[2 * k];
  const int k2 = k + (ms + 1);
  const bool move_k2 = F.prop_mode ? (F.sm.select[k] != 0) : true;
  const int stay = move_k2 ? k : k2, go = move_k2 ? k2 : k;
  const LabelStats& st = F.st;
  const LabelStats& h = F.sst;
  st.count[k] = h.count[stay];
  st.count[s] = h.count[go];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    st.S[3 * k + c] = h.S[3 * stay + c];
    st.S[3 * s + c] = h.S[3 * go + c];
    st.Q[3 * k + c] = h.Q[3 * stay + c];
    st.Q[3 * s + c] = h.Q[3 * go + c];
  }
#pragma unroll
  for (int c = 0; c < 5; c++) {
    st.X[5 * k + c] = h.X[5 * stay + c];
    st.X[5 * s + c] = h.X[5 * go + c];
  }
}

__global__ void kb_add_max(Batch b) {
  int f = threadIdx.x;
  if (f >= b.nf) return;
  const FrameDesc& F = b.fr[f];
  const int ms = *F.max_spix;
  if (ms > F.nbuf / 2) return;
  *F.max_spix = ms + min(*F.total, F.nbuf - 1 - ms);
}

// The three BFS kernels write disjoint pixels (different labels): the few
// long workgroup searches run on two side streams beside the wave searches
// (all on the main stream when the batch has no side streams).
// Expects the halves map to equal the labels: the statistics pass with
// bounding boxes that precedes every split writes it.
void b_split(const Batch& b, int cur, int direction, hipStream_t s, const SideStream& side) {
  kb_sm_clear<<<b.sm_blocks, kBlock, 0, s>>>(b);
  kb_reset_work<<<1, 64, 0, s>>>(b);
  kb_split_list<<<b.lab_blocks, kBlock, 0, s>>>(b);
  if (side.st) {
    (void)hipEventRecord(side.fork, s);
    (void)hipStreamWaitEvent(side.st, side.fork, 0);
    (void)hipStreamWaitEvent(side.st2, side.fork, 0);
  }
  kb_bfs_big<<<64, 512, 0, side.st ? side.st2 : s>>>(b, cur, direction);
  kb_bfs_block<<<256, 256, 0, side.st ? side.st : s>>>(b, cur, direction);
  kb_bfs_wave<<<2048, 256, 0, s>>>(b, cur, direction);
  if (side.st) {
    (void)hipEventRecord(side.join, side.st);
    (void)hipEventRecord(side.join2, side.st2);
    (void)hipStreamWaitEvent(s, side.join, 0);
    (void)hipStreamWaitEvent(s, side.join2, 0);
  }
  kb_stats_accum<false, true><<<b.st_blocks, 256, 0, s>>>(b, cur, 1);
  kb_split_eval<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_scan<<<b.nf, 1024, 0, s>>>(b, 0);
  kb_split_assign<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_split_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_split_stats<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_add_max<<<1, 64, 0, s>>>(b);
}

// ---------------------------------------------------------------------------
// merge (split_merge_orig.cu:79-199, split_merge_prop.cu:87-181)
// ---------------------------------------------------------------------------
// Merge candidates of border pixels (a pixel on the frame edge or with a
// 4-neighbour of another label), the border test done inline.
__global__ void kb_merge_cand(Batch b, int cur, int direction) {
  PIX_PROLOGUE
  if (t >= F.npix) return;
  const int W = F.W, H = F.H;
  const int* seg = F.seg[cur];
  const int prop_mode = F.prop_mode;
  int x = t % W, y = t / W;
  int C = seg[t], Wn = -1;
  if (y >= 1 && x >= 1 && y < H - 1 && x < W - 1 && C == seg[t - W] && C == seg[t + W] &&
      C == seg[t + 1] && C == seg[t - 1])
    return;
  if (direction == 1) {
    bool ok = prop_mode ? (y > 1 && y < H - 2) : (y >= 0 && y < H - 1);
    if (ok) Wn = seg[t + W];
  } else {
    bool ok = prop_mode ? (x > 1 && x < W - 2) : (x > 0 && x < W);
    if (ok) Wn = seg[t - 1];
  }
  if (Wn >= 0 && C != Wn) {
    if (prop_mode && (F.P.prop[C] == F.P.prop[Wn])) return;
    atomicMax(F.sm.pairs + 2 * C + 1, Wn);
  }
}

__global__ void kb_merge_step0(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf || F.P.valid[k] == 0) return;
  SmState sm = F.sm;
  int fk = sm.pairs[2 * k + 1];
  if (fk < 0) return;
  const NigConsts& nc = F.nc;
  int count_f = F.P.count[fk], count_k = F.P.count[k];
  int count_fk = count_f + count_k;
  sm.count_f[k] = count_fk;
  for (int c = 0; c < 3; c++) {
    double Qk = fx_to_double(F.sst.Q[3 * k + c]), Qf = fx_to_double(F.sst.Q[3 * fk + c]);
    double Sk = fx_to_double(F.st.S[3 * k + c]), Sf = fx_to_double(F.st.S[3 * fk + c]);
    double bn = nc.b0 + 0.5 * (Qk - (Sk * Sk) / (double)count_k);
    double bf = nc.b0 + 0.5 * ((Qk + Qf) - ((Sf + Sk) * (Sf + Sk)) / (double)count_fk);
    if (bn < 0) bn = 0.1;
    if (bf < 0) bf = 0.1;
    sm.b_n[3 * k + c] = bn;
    sm.b_n_f[3 * k + c] = bf;
  }
  sm.has_bn[k] = 1;
}

__global__ void kb_merge_marg(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf || F.P.valid[k] == 0) return;
  F.sm.marg[k] = nig_marg(F.nc, (double)F.P.count[k], F.sm.b_n + 3 * k);
}

__global__ void kb_merge_eval(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf || F.P.valid[k] == 0) return;
  SmState sm = F.sm;
  int fk = sm.pairs[2 * k + 1];
  if (fk < 0 || F.P.valid[fk] == 0) return;
  int count_k = F.P.count[k], count_fk = sm.count_f[k];
  if (count_k < 1 || count_fk < 1) return;
  double mfk = nig_marg(F.nc, (double)count_fk, sm.b_n_f + 3 * k);
  double h = mfk - (F.nc.log_alpha + sm.marg[k] + sm.marg[fk]) + (double)F.merge_alpha;
  sm.hasting[k] = h;
  if (h > -2.0) atomicMax(sm.winner + fk, k);
}

__global__ void kb_merge_flag(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  SmState sm = F.sm;
  int w = sm.winner[k];
  if (w >= 0) {
    sm.merge[w] = 1;
    sm.pairs[2 * k] = w;
  }
  sm.valid0[k] = F.P.valid[k];
  sm.pc0[k] = F.P.prior_count[k];
  sm.rank[k] = sm.pairs[2 * k + 1];  // target snapshot
}

__global__ void kb_merge_remove(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  SmState sm = F.sm;
  Params P = F.P;
  const int* tgt0 = sm.rank;
  int s = tgt0[k];
  if (s < 0) return;
  int is_cycle = tgt0[s] == k;
  if (sm.valid0[k] == 0 || sm.valid0[s] == 0) return;
  int mk = sm.merge[k], msg = sm.merge[s];
  if (!F.prop_mode) {
    if (mk && (!msg || (is_cycle && k < s))) {
      sm.remove[k] = 1;
      P.valid[k] = 0;
      P.prior_count[s] = sm.pc0[k] + sm.pc0[s];
    }
  } else {
    int is_prop_k = P.prop[k];
    int cond_c = is_cycle && is_prop_k;
    if (mk && (!msg || cond_c)) {
      int victim = k, keeper = s;
      if (is_prop_k) {
        sm.pairs[2 * s + 1] = k;
        victim = s;
        keeper = k;
      }
      sm.remove[victim] = 1;
      P.valid[victim] = 0;
      P.prior_count[keeper] = sm.pc0[victim] + sm.pc0[keeper];
    }
  }
}

__global__ void kb_merge_apply(Batch b, int cur) {
  PIX_PROLOGUE
  if (t >= F.npix) return;
  int* seg = F.seg[cur];
  int k = seg[t];
  if (k >= 0 && k < F.nbuf && F.sm.remove[k]) seg[t] = F.sm.pairs[2 * k + 1];
}

// Label statistics after a merge without a recount: the merge's own pass
// gathered every live label's exact integer sums (into the split/merge
// slots), and kb_merge_apply moves the pixels of a removed label k to
// pairs[2k + 1] in one hop, so each label's sums are added to its
// destination (after kb_stats_clear).
__global__ void kb_merge_stats(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const LabelStats& h = F.sst;
  const int n = h.count[k];
  if (n == 0) return;
  const int x = F.sm.remove[k] ? F.sm.pairs[2 * k + 1] : k;
  const LabelStats& st = F.st;
  atomicAdd(st.count + x, n);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    atomic_add_i64(st.S + 3 * x + c, h.S[3 * k + c]);
    atomic_add_i64(st.Q + 3 * x + c, h.Q[3 * k + c]);
  }
#pragma unroll
  for (int c = 0; c < 5; c++) atomic_add_i64(st.X + 5 * x + c, h.X[5 * k + c]);
}

// The merge's statistics of the current labels without a pixel pass: the
// live statistics (kept exact by the sweeps, the split and the merge) moved
// along the relabel map of kb_relabel_apply (a label L goes to target[L]
// when that is set, in one hop; propagated frames only).
__global__ void kb_merge_gather(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const LabelStats& st = F.st;
  const int n = st.count[k];
  if (n == 0) return;
  const int x = (F.prop_mode && F.target[k] >= 0) ? F.target[k] : k;
  const LabelStats& h = F.sst;
  atomicAdd(h.count + x, n);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    atomic_add_i64(h.S + 3 * x + c, st.S[3 * k + c]);
    atomic_add_i64(h.Q + 3 * x + c, st.Q[3 * k + c]);
  }
#pragma unroll
  for (int c = 0; c < 5; c++) atomic_add_i64(h.X + 5 * x + c, st.X[5 * k + c]);
}

void b_merge(const Batch& b, int cur, int direction, hipStream_t s) {
  kb_sm_clear<<<b.sm_blocks, kBlock, 0, s>>>(b);
  kb_merge_cand<<<b.pix_blocks, kBlock, 0, s>>>(b, cur, direction);
  kb_merge_gather<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_step0<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_marg<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_eval<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_flag<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_remove<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_stats_clear<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
  kb_merge_stats<<<b.lab_blocks, kBlock, 0, s>>>(b);
}

// ---------------------------------------------------------------------------
// relabel (relabel.cu:85-409), rules R7, R8.  The only candidates the
// reference's rule can pick for a living label are the previous-frame ids
// that are dead now, and the label itself; the search scans exactly those.
// ---------------------------------------------------------------------------
__global__ void kb_relabel_prep(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  F.alive[k] = 0;
  F.target[k] = -1;
  F.want_new[k] = 0;
  F.src_of[k] = -1;
}

__global__ void kb_mark_alive(Batch b, int cur) {
  PIX_PROLOGUE
  if (t >= F.npix) return;
  int k = F.seg[cur][t];
  if (k >= 0 && k < F.nbuf) F.alive[k] = 1;
}

__global__ void kb_relabel_lists(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const int nspix = *F.max_spix + 1;
  if (k < F.nspix_prev && !F.alive[k]) F.dead[atomicAdd(F.counter, 1)] = k;
  if (k < nspix && F.alive[k]) {
    int pos = atomicAdd(b.work_n, 1);
    if (pos < b.work_cap) b.work[pos] = (f << 24) | k;
  }
}

// Persistent waves: one living label per wave at a time.  The loop has no
// lane-dependent branch: every lane takes part in the atomic (only lane 0
// adds 1) and the item index is made wave-uniform with readfirstlane, so the
// loop exit is a scalar branch.  (With `if (lane == 0) it = atomicAdd(..)`
// plus a lane-0 store at the end of the body, the compiler threads the
// lane-0 paths across the back-edge into a loop nest in which lanes 1..63
// re-read a reset index from lane 0 and never leave.)
__global__ void __launch_bounds__(256) kb_relabel_search(Batch b) {
  const int lane = threadIdx.x & 63;
  const int n_items = __builtin_amdgcn_readfirstlane(min(b.work_n[0], b.work_cap));
  const int gw = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
  const int nwaves = gridDim.x * 4;
  for (int it = gw; it < n_items; it += nwaves) {
    const int wv = b.work[it];
    const int f = wv >> 24, s = wv & 0xffffff;
    const FrameDesc& F = b.fr[f];
    const float* mu = F.P.mu_app + 3 * s;
    const float m0 = mu[0], m1 = mu[1], m2 = mu[2];
    const float* pm = F.prev_prior_mu_app;
    const int ndead = *F.counter;
    float best = 10000.0f;
    int cand = -1;
    float ss = 0.0f;
    // candidates: the dead previous ids, then the label itself
    for (int j = lane; j <= ndead; j += 64) {
      int c;
      if (j < ndead) {
        c = F.dead[j];
      } else {
        if (s >= F.nspix_prev) continue;
        c = s;
      }
      float dx = m0 - pm[3 * c];
      float dy = m1 - pm[3 * c + 1];
      float dz = m2 - pm[3 * c + 2];
      float delta = (dx * dx + dy * dy + dz * dz);
      if (c == s) ss = delta;
      if (delta < best || (delta == best && cand >= 0 && c < cand)) {
        best = delta;
        cand = c;
      }
    }
    for (int d = 32; d >= 1; d >>= 1) {
      float b2 = __shfl_xor(best, d);
      int c2 = __shfl_xor(cand, d);
      bool take = (c2 >= 0) && (cand < 0 || b2 < best || (b2 == best && c2 < cand));
      if (take) {
        best = b2;
        cand = c2;
      }
    }
    for (int d = 32; d >= 1; d >>= 1) ss = fmaxf(ss, __shfl_xor(ss, d));
    // every lane holds the reduced values; all lanes store the same words
    const int nspix = *F.max_spix + 1;
    const bool revive = (cand >= 0) && (best < F.thresh_relabel) && (cand < nspix) && (cand != s);
    F.target[s] = revive ? cand : -1;
    F.want_new[s] = (!revive && ss > F.thresh_new) ? 1 : 0;
  }
}

__global__ void kb_relabel_assign(Batch b) {
  LAB_PROLOGUE
  const int nspix = *F.max_spix + 1;
  if (k >= nspix || k >= F.nbuf) return;
  if (F.want_new[k]) {
    int id = nspix + F.rank[k];
    F.target[k] = id < F.nbuf ? id : -1;
  }
  int t = F.target[k];
  if (t >= 0) atomicMax(F.src_of + t, k);
}

__global__ void kb_relabel_params(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  const int s = F.src_of[k];
  if (s < 0) return;
  Params P = F.P;
  P.count[k] = P.count[s];
  P.icount[k] = P.icount[s];
  if (k >= F.nspix_prev) {
    P.prop[k] = 0;
    P.prior_count[k] = P.prior_count[s];
    for (int c = 0; c < 3; c++) {
      P.prior_sigma_shape[3 * k + c] = P.prior_sigma_shape[3 * s + c];
      P.sample_sigma_shape[3 * k + c] = P.sample_sigma_shape[3 * s + c];
      P.prior_mu_app[3 * k + c] = P.prior_mu_app[3 * s + c];
      P.prior_icov[3 * k + c] = P.prior_icov[3 * s + c];
    }
  }
}

__global__ void kb_relabel_valid(Batch b) {
  LAB_PROLOGUE
  if (k > *F.max_spix || k >= F.nbuf) return;
  int t = F.target[k];
  if (t >= 0) {
    F.P.valid[k] = 0;
    F.P.valid[t] = 1;
  }
}

__global__ void kb_relabel_apply(Batch b, int cur) {
  PIX_PROLOGUE
  if (t >= F.npix) return;
  int* seg = F.seg[cur];
  int L = seg[t];
  if (L >= 0 && L < F.nbuf) {
    int tg = F.target[L];
    if (tg >= 0) seg[t] = tg;
  }
}

__global__ void kb_relabel_max(Batch b) {
  int f = threadIdx.x;
  if (f >= b.nf) return;
  const FrameDesc& F = b.fr[f];
  int m = *F.max_spix + *F.total;
  if (m > F.nbuf - 1) m = F.nbuf - 1 > *F.max_spix ? F.nbuf - 1 : *F.max_spix;
  *F.max_spix = m;
}

void b_relabel(const Batch& b, int cur, hipStream_t s) {
  kb_relabel_prep<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_mark_alive<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_reset_work<<<1, 64, 0, s>>>(b);
  kb_relabel_lists<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_search<<<512, 256, 0, s>>>(b);
  kb_scan<<<b.nf, 1024, 0, s>>>(b, 1);
  kb_relabel_assign<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_params<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_valid<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_relabel_max<<<1, 64, 0, s>>>(b);
}

}  // namespace stsp
