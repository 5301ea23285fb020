// Batched refinement-loop kernels: every launch covers all frames of a batch
// (see batch.h).  Semantics per frame are those of the reference kernels cited
// at each function.  Where the reference depends on atomic arrival order the
// rule here is order-independent:
//   - colour/position sums are exact int64 fixed-point additions;
//   - the merge winner of a target f is the largest accepting k (atomicMax);
//   - new split / relabel ids are max_spix+1+rank in ascending label order;
//   - remove steps read valid / prior_count / pairs from a snapshot;
//   - the relabel candidate is the lexicographic minimum of (distance, id).
#include "kb_dev.h"

namespace stsp {

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

// Statistics of the live labels (limit nbuf) into st, or with MOVED those
// of the parts that a split search moved to new labels (the halves map's
// labels above max_spix) into sst.
//
// A workgroup covers a kStatTile x kStatTile tile: wave w its rows 8w..8w+7,
// pixel slot j of lane l column l&31 of row 8w + 2j + (l>>5) (each slot's
// loads cover two whole 128-byte row segments).  For each distinct
// label of a wave (usually one to three) the lanes sum their pixels of that
// label, DPP folds the sums into lane 63, and lane 63 issues one set of
// atomics.  Position moments are summed relative to the tile origin in int32
// and widened exactly: sum x = n*x0 + sum dx, sum x^2 = n*x0^2 + 2*x0*sum dx
// + sum dx^2, sum xy = n*x0*y0 + x0*sum dy + y0*sum dx + sum dx*dy.
// SQ: also gather the colour squares (split and merge read them; the sweeps
// keep them current with the other sums).
// BBOX: also the bounding boxes, the halves map (a copy of the labels) and,
// for propagated frames, the count of pixels whose shifted label is invalid
// (ninvalid): a full pass at the first iteration when it splits (the pass
// before a later split, whose sums are current, is kb_bbox_accum).
template <bool BBOX, bool SQ, bool MOVED = false>
__global__ void __launch_bounds__(256) kb_stats_accum(Batch b, int cur) {
  static_assert(!MOVED || (SQ && !BBOX), "the moved parts take every sum");
  const int f = frame_of(b.off.st, b.nf, blockIdx.x);
  const FrameDesc& F = b.fr[f];
  const int W = F.W, H = F.H;
  const int tile = blockIdx.x - b.off.st[f];
  const int tiles_x = (W + kStatTile - 1) / kStatTile;
  const int tx0 = (tile % tiles_x) * kStatTile, ty0 = (tile / tiles_x) * kStatTile;
  const int lane = threadIdx.x & 63;
  const int dx = lane & 31, wrow = 8 * (threadIdx.x >> 6) + (lane >> 5);
  const int x = tx0 + dx;
  const int* labs = MOVED ? F.halves : F.seg[cur];
  LabelStats st = MOVED ? F.sst : F.st;
  const int nlab = MOVED ? F.nsm : F.nbuf;
  const int lo = MOVED ? *F.max_spix + 1 : 0;  // the moved parts: k + max_spix + 1
  const int* shifted = ((BBOX || MOVED) && F.prop_mode) ? F.shifted : nullptr;
  int k[4];
  float col[4][3];
  int inv[4];
  unsigned pend = 0;
  // the labels of the lane's four pixels are loaded before any is used
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int y = ty0 + wrow + 2 * j;
    k[j] = (x < W && y < H) ? labs[y * W + x] : -1;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int y = ty0 + wrow + 2 * j;
    const int t = y * W + x;
    const int v = k[j];
    inv[j] = 0;
    col[j][0] = col[j][1] = col[j][2] = 0.f;
    // the pass before a split also starts the halves map (a copy of the
    // labels; the split search then writes the moving parts)
    if (BBOX && x < W && y < H) F.halves[t] = v;
    if (v >= lo && v < nlab) {
      pend |= 1u << j;
      const float* p = F.img + 3 * t;
      col[j][0] = p[0];
      col[j][1] = p[1];
      col[j][2] = p[2];
      if (shifted) inv[j] = shifted[t] < 0;
    } else {
      k[j] = -1;
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
    int mny = 0x7fffffff, mxy = -1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (((pend >> j) & 1u) && k[j] == u) {
        pend &= ~(1u << j);
        const int dy = wrow + 2 * j;
        cnt += 1;
        ninv += inv[j];
        sdx += dx;
        sdy += dy;
        sxx += dx * dx;
        sxy += dx * dy;
        syy += dy * dy;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          S[c] += fx_color(col[j][c]);
          if (SQ) Q[c] += fx_square(col[j][c]);
        }
        if (BBOX) {
          mny = min(mny, ty0 + dy);
          mxy = max(mxy, ty0 + dy);
        }
      }
    }
    int mnx = cnt ? x : 0x7fffffff, mxx = cnt ? x : -1;
    if (shifted) ninv = wave_sum_to63(ninv);
    cnt = wave_sum_to63(cnt);
    sdx = wave_sum_to63(sdx);
    sdy = wave_sum_to63(sdy);
    sxx = wave_sum_to63(sxx);
    sxy = wave_sum_to63(sxy);
    syy = wave_sum_to63(syy);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      S[c] = wave_sum64_to63(S[c]);
      if (SQ) Q[c] = wave_sum64_to63(Q[c]);
    }
    if (BBOX) {
      mnx = wave_min_to63(mnx);
      mxx = wave_max_to63(mxx);
      mny = wave_min_to63(mny);
      mxy = wave_max_to63(mxy);
    }
    if (lane == 63) {
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
  param_update_one(F, k, summary_mode);
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

// The pass before a split: bounding boxes and invalid counts of the live
// labels, and the halves map (a copy of the labels).  A workgroup covers a
// kStatTile x kStatTile tile, wave w its rows 8w..8w+7; pixel slot j of a
// lane is column lane & 31 of row 8w + 2j + (lane >> 5), so every load and
// store of a slot covers two whole 128-byte row segments.  Per distinct
// label of a wave the extents come from ballots of its pixels, in scalar
// registers: the columns from the OR of the slots' half-wave masks, the rows
// from which half-wave masks are non-empty.
__global__ void __launch_bounds__(256) kb_bbox_accum(Batch b, int cur) {
  static_assert(kStatTile == 32, "a half-wave covers one tile row");
  const int f = frame_of(b.off.st, b.nf, blockIdx.x);
  const FrameDesc& F = b.fr[f];
  const int W = F.W, H = F.H;
  const int tile = blockIdx.x - b.off.st[f];
  const int tiles_x = (W + kStatTile - 1) / kStatTile;
  const int tx0 = (tile % tiles_x) * kStatTile, ty0 = (tile / tiles_x) * kStatTile;
  const int lane = threadIdx.x & 63;
  const int wrow = 8 * (threadIdx.x >> 6);  // the wave's first row of the tile
  const int x = tx0 + (lane & 31), y0 = ty0 + wrow + (lane >> 5);
  const int* labs = F.seg[cur];
  const int* shifted = F.prop_mode ? F.shifted : nullptr;
  // all loads of the lane's four pixels go out before the first is used
  int k[4], sh[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int y = y0 + 2 * j;
    const bool in = x < W && y < H;
    const int t = in ? y * W + x : 0;
    k[j] = in ? labs[t] : -1;
    sh[j] = (in && shifted) ? shifted[t] : 0;
  }
  unsigned pend = 0, invm = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int y = y0 + 2 * j;
    if (x < W && y < H) F.halves[y * W + x] = k[j];
    if (k[j] >= 0 && k[j] < F.nbuf) {
      pend |= 1u << j;
      if (sh[j] < 0) invm |= 1u << j;
    } else {
      k[j] = -1;
    }
  }
  while (true) {
    const uint64_t act = __ballot(pend != 0);
    if (act == 0) break;
    const int leader = __builtin_ctzll(act);
    const int first = (pend & 1u) ? k[0] : (pend & 2u) ? k[1] : (pend & 4u) ? k[2] : k[3];
    const int u = __builtin_amdgcn_readlane(first, leader);
    unsigned mine = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (((pend >> j) & 1u) && k[j] == u) mine |= 1u << j;
    pend &= ~mine;
    uint32_t cols = 0, rows = 0;  // rows: bit 2j + h of the wave's 8 rows
    int ninv = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t m = __ballot((mine >> j) & 1u);
      const uint32_t lo = (uint32_t)m, hi = (uint32_t)(m >> 32);
      cols |= lo | hi;
      rows |= (lo ? 1u : 0u) << (2 * j);
      rows |= (hi ? 1u : 0u) << (2 * j + 1);
      if (shifted) ninv += __popcll(__ballot(((mine & invm) >> j) & 1u));
    }
    if (lane == 0) {
      int* bb = F.st.bbox + 4 * u;
      atomicMin(bb, tx0 + __builtin_ctz(cols));
      atomicMax(bb + 1, tx0 + 31 - __builtin_clz(cols));
      atomicMin(bb + 2, ty0 + wrow + __builtin_ctz(rows));
      atomicMax(bb + 3, ty0 + wrow + 31 - __builtin_clz(rows));
      if (ninv) atomicAdd(F.st.ninvalid + u, ninv);
    }
  }
}

// parameters from the statistics as they stand (kept current by the sweeps)
void b_param_only(const Batch& b, hipStream_t s) {
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

void b_update_params(const Batch& b, int cur, bool bbox, hipStream_t s) {
  kb_stats_clear<<<b.lab_blocks, kBlock, 0, s>>>(b, bbox ? 1 : 0);
  if (bbox)
    kb_stats_accum<true, true><<<b.st_blocks, 256, 0, s>>>(b, cur);
  else
    kb_stats_accum<false, true><<<b.st_blocks, 256, 0, s>>>(b, cur);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

__global__ void kb_bbox_clear(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  F.st.ninvalid[k] = 0;
  int* bb = F.st.bbox + 4 * k;
  bb[0] = 0x7fffffff;
  bb[1] = -1;
  bb[2] = 0x7fffffff;
  bb[3] = -1;
}

// Before a split: the label sums are current (the sweeps move them, split and
// merge combine them), so the pass gathers only the bounding boxes and the
// invalid counts (which the sweeps do not keep) and starts the halves map;
// then the parameters are updated.
void b_prepare_split(const Batch& b, int cur, hipStream_t s) {
  kb_bbox_clear<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_bbox_accum<<<b.st_blocks, 256, 0, s>>>(b, cur);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
}

void b_param_summary(const Batch& b, int cur, hipStream_t s) {
  kb_stats_clear<<<b.lab_blocks, kBlock, 0, s>>>(b, 0);
  kb_stats_accum<false, false><<<b.st_blocks, 256, 0, s>>>(b, cur);
  kb_param_update<<<b.lab_blocks, kBlock, 0, s>>>(b, 1);
}

void b_store_sample(const Batch& b, hipStream_t s) {
  kb_store_sample<<<b.lab_blocks, kBlock, 0, s>>>(b);
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
// (split accept -> sm.rank) or (relabel want_new -> rank); total -> F.total.
// Tiles of 4096 entries (4 consecutive per thread), the next tile's loads in
// flight while one is scanned: a thread's 4 sums, a wave scan of the thread
// totals, then the 16 wave totals through LDS.
__global__ void __launch_bounds__(1024) kb_scan(Batch b, int which) {
  const FrameDesc& F = b.fr[blockIdx.x];
  const int* __restrict__ in = which == 0 ? F.sm.accept : F.want_new;
  int* __restrict__ out = which == 0 ? F.sm.rank : F.rank;
  const int n = F.nbuf;
  constexpr int kTile = 4096;
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int v[4], nx[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int i = 4 * tid + j;
    v[j] = i < n ? in[i] : 0;
  }
  int carry = 0;
  for (int base = 0; base < n; base += kTile) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = base + kTile + 4 * tid + j;
      nx[j] = i < n ? in[i] : 0;
    }
    const int s0 = v[0], s1 = s0 + v[1], s2 = s1 + v[2], s3 = s2 + v[3];
    int x = s3;  // inclusive scan of the thread totals over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int wpre = 0, tile = 0;
#pragma unroll
    for (int w = 0; w < 16; w++) {
      const int t = wsum[w];
      wpre += w < wv ? t : 0;
      tile += t;
    }
    const int ex = carry + wpre + x - s3;
    const int pre[4] = {0, s0, s1, s2};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = base + 4 * tid + j;
      if (i < n) out[i] = ex + pre[j];
    }
    carry += tile;
    __syncthreads();  // wsum is rewritten by the next tile
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = nx[j];
  }
  if (tid == 0) *F.total = carry;
}


// ---------------------------------------------------------------------------
// split (split_merge_orig.cu:202-420, split_merge_prop.cu:184-395)
// ---------------------------------------------------------------------------
__global__ void kb_sm_clear(Batch b) {
  // block 0 also clears the work-list counters and each frame's counter
  if (blockIdx.x == 0) {
    if (threadIdx.x < kWorkCounters) b.work_n[threadIdx.x] = 0;
    if (threadIdx.x < b.nf) b.fr[threadIdx.x].counter[0] = 0;
  }
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


// ---------------------------------------------------------------------------
// merge (split_merge_orig.cu:79-199, split_merge_prop.cu:87-181)
// ---------------------------------------------------------------------------
// Merge candidates of border pixels (a pixel on the frame edge or with a
// 4-neighbour of another label), the border test done inline.
__global__ void kb_merge_cand(Batch b, int cur, int direction) {
  PIX_PROLOGUE
  const int W = F.W, H = F.H;
  const int* __restrict__ seg = F.seg[cur];
  // (x, y) of pixel t0, then stepped by kBlock pixels (no division per
  // pixel); the labels of all four pixels and their neighbours are loaded
  // before any is used
  int xs[kPixPerThread], ys[kPixPerThread], C[kPixPerThread], Wn[kPixPerThread];
  bool border[kPixPerThread];
  {
    int y = t0 / W, x = t0 - y * W;
#pragma unroll
    for (int j = 0; j < kPixPerThread; j++) {
      if (j > 0) {
        x += kBlock;
        while (x >= W) {
          x -= W;
          y++;
        }
      }
      xs[j] = x;
      ys[j] = y;
    }
  }
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    const int t = t0 + j * kBlock, x = xs[j], y = ys[j];
    const bool valid = t < F.npix;
    const int c = valid ? seg[t] : -1;
    const int N = valid && y > 0 ? seg[t - W] : c, S = valid && y < H - 1 ? seg[t + W] : c;
    const int Wl = valid && x > 0 ? seg[t - 1] : c, E = valid && x < W - 1 ? seg[t + 1] : c;
    const bool inner = y >= 1 && x >= 1 && y < H - 1 && x < W - 1 && c == N && c == S &&
                       c == E && c == Wl;
    int wn = -1;
    if (valid && !inner) {
      if (direction == 1) {
        const bool ok = F.prop_mode ? (y > 1 && y < H - 2) : (y >= 0 && y < H - 1);
        if (ok) wn = S;
      } else {
        const bool ok = F.prop_mode ? (x > 1 && x < W - 2) : (x > 0 && x < W);
        if (ok) wn = Wl;
      }
    }
    C[j] = c;
    Wn[j] = wn >= 0 && c != wn ? wn : -1;
    border[j] = Wn[j] >= 0;
  }
  // prop mode: only pairs of labels that are not both propagated or both new
  bool take[kPixPerThread];
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++)
    take[j] = border[j] && !(F.prop_mode && F.P.prop[C[j]] == F.P.prop[Wn[j]]);
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++)
    if (take[j]) atomicMax(F.sm.pairs + 2 * C[j] + 1, Wn[j]);
}

__global__ void kb_merge_step0(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf || F.P.valid[k] == 0) return;
  SmState sm = F.sm;
  int fk = sm.pairs[2 * k + 1];
  // the marginal of the label alone (b_n stays cleared without a candidate)
  if (fk < 0) {
    sm.marg[k] = nig_marg(F.nc, (double)F.P.count[k], sm.b_n + 3 * k);
    return;
  }
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
  sm.marg[k] = nig_marg(nc, (double)count_k, sm.b_n + 3 * k);
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
  // the live statistics are cleared for kb_merge_stats (nothing reads them
  // between here and there)
  const LabelStats& st = F.st;
  st.count[k] = 0;
  st.ninvalid[k] = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    st.S[3 * k + c] = 0;
    st.Q[3 * k + c] = 0;
  }
#pragma unroll
  for (int c = 0; c < 5; c++) st.X[5 * k + c] = 0;
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
  int* seg = F.seg[cur];
  int k[kPixPerThread];
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    const int t = t0 + j * kBlock;
    k[j] = t < F.npix ? seg[t] : -1;
  }
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++)
    if (k[j] >= 0 && k[j] < F.nbuf && F.sm.remove[k[j]])
      seg[t0 + j * kBlock] = F.sm.pairs[2 * k[j] + 1];
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
  kb_merge_eval<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_flag<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_remove<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_merge_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_merge_stats<<<b.lab_blocks, kBlock, 0, s>>>(b);
}

// ---------------------------------------------------------------------------
// relabel (relabel.cu:85-409), rules R7, R8.  The only candidates the
// reference's rule can pick for a living label are the previous-frame ids
// that are dead now, and the label itself; the search scans exactly those.
// ---------------------------------------------------------------------------
__global__ void kb_relabel_prep(Batch b) {
  // block 0 also clears the work-list counters and each frame's counter
  if (blockIdx.x == 0) {
    if (threadIdx.x < kWorkCounters) b.work_n[threadIdx.x] = 0;
    if (threadIdx.x < b.nf) b.fr[threadIdx.x].counter[0] = 0;
  }
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  F.alive[k] = 0;
  F.target[k] = -1;
  F.want_new[k] = 0;
  F.src_of[k] = -1;
}

__global__ void kb_mark_alive(Batch b, int cur) {
  PIX_PROLOGUE
  int k[kPixPerThread];
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    const int t = t0 + j * kBlock;
    k[j] = t < F.npix ? F.seg[cur][t] : -1;
  }
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++)
    if (k[j] >= 0 && k[j] < F.nbuf) F.alive[k[j]] = 1;
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

// also the valid flags: a label with a target gives up its id for it
// (nothing else in this pass reads them)
__global__ void kb_relabel_params(Batch b) {
  LAB_PROLOGUE
  if (k >= F.nbuf) return;
  if (k <= *F.max_spix) {
    const int t = F.target[k];
    if (t >= 0) {
      F.P.valid[k] = 0;
      F.P.valid[t] = 1;
    }
  }
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


__global__ void kb_relabel_apply(Batch b, int cur) {
  // block 0 also moves each frame's max_spix past the new ids (no pixel of
  // this pass reads it)
  if (blockIdx.x == 0 && threadIdx.x < b.nf) {
    const FrameDesc& G = b.fr[threadIdx.x];
    int m = *G.max_spix + *G.total;
    if (m > G.nbuf - 1) m = G.nbuf - 1 > *G.max_spix ? G.nbuf - 1 : *G.max_spix;
    *G.max_spix = m;
  }
  PIX_PROLOGUE
  int* seg = F.seg[cur];
  int L[kPixPerThread];
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    const int t = t0 + j * kBlock;
    L[j] = t < F.npix ? seg[t] : -1;
  }
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    if (L[j] < 0 || L[j] >= F.nbuf) continue;
    const int tg = F.target[L[j]];
    if (tg >= 0) seg[t0 + j * kBlock] = tg;
  }
}


// launches shared with the split step (k_split.hip)
void b_sm_begin(const Batch& b, hipStream_t s) {
  kb_sm_clear<<<b.sm_blocks, kBlock, 0, s>>>(b);
}

void b_moved_stats(const Batch& b, int cur, hipStream_t s) {
  kb_stats_accum<false, true, true><<<b.st_blocks, 256, 0, s>>>(b, cur);
}

void b_scan(const Batch& b, int which, hipStream_t s) {
  kb_scan<<<b.nf, 1024, 0, s>>>(b, which);
}

void b_relabel(const Batch& b, int cur, hipStream_t s) {
  kb_relabel_prep<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_mark_alive<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_relabel_lists<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_search<<<512, 256, 0, s>>>(b);
  kb_scan<<<b.nf, 1024, 0, s>>>(b, 1);
  kb_relabel_assign<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_params<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_relabel_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
}

}  // namespace stsp
