// Split and merge moves (reference split_merge_orig.cu, split_merge.cu:358-1080,
// split_merge_prop.cu).  The Hastings ratios are evaluated in FP64 from the
// deterministic int64 statistics; every first-arrival / atomicAdd-order choice
// of the reference is replaced by an order-independent rule:
//   - merge winner of a target f is the largest accepting k  (atomicMax)
//   - new split ids are max_spix+1+rank of k among accepted k (prefix sum)
//   - remove reads valid / prior_count / pairs from a snapshot.
#include "kernels.h"

namespace stsp {

// NIG marginal (split_merge_prop.cu:2975-3019); lgamma terms cancel exactly.
__device__ __forceinline__ double nig_marginal(const NigConsts& c, double n, const double* bn) {
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

__global__ void k_sm_clear(SmState sm, LabelStats sst, int nsm, int nbuf) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsm) return;
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
  if (k < nbuf) {
    sm.pairs[2 * k] = -1;
    sm.pairs[2 * k + 1] = -1;
  }
}

void launch_sm_clear(SmState sm, LabelStats sst, int nsm, int nbuf, hipStream_t s) {
  k_sm_clear<<<div_up(nsm, kBlock), kBlock, 0, s>>>(sm, sst, nsm, nbuf);
}

// ---------------------------------------------------------------------------
// split proposal: seeds, per-label BFS, halves
// ---------------------------------------------------------------------------
// Seed pixel of label k for one direction (init_split(_p),
// split_merge.cu:803-845): (int)mu +- direction, kept only if it carries k.
__device__ __forceinline__ int split_seed(const Params& P, const int* __restrict__ seg, int k,
                                          int dir, int W, int H) {
  int x, y;
  if (dir == 1 || dir == -1) {
    x = (int)(P.mu_shape[2 * k]) + dir;
    y = (int)(P.mu_shape[2 * k + 1]);
  } else {
    x = (int)(P.mu_shape[2 * k]);
    y = (int)(P.mu_shape[2 * k + 1]) + dir;
  }
  int ind = y * W + x;
  if (ind < 0 || ind > W * H - 1) return -1;
  return seg[ind] == k ? ind : -1;
}

__global__ void k_copy_i32(const int* __restrict__ a, int* __restrict__ b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Two-seed split proposal of one label (calc_split_candidate, split_merge.cu:
// 757-800 run to convergence from both seeds, then calc_seg_split 847-860).
// One workgroup per label runs the level-synchronous 4-connected BFS inside
// the label's bounding box with both distance maps in LDS; a label whose box
// exceeds the LDS budget runs the same levels on the global d1/d2 maps.
// Pixels with d1 > d2 (unreached = -1) go to half k + max_spix + 1.
constexpr int kBfsCap = 8192;

__global__ void __launch_bounds__(256)
    k_bfs_box(const int* __restrict__ seg, Params P, const int* __restrict__ bbox, int nbuf,
              const int* __restrict__ max_spix, int direction, int W, int H, int* d1g, int* d2g,
              int* halves) {
  const int k = blockIdx.x;
  const int ms = *max_spix;
  if (k >= nbuf || ms > nbuf / 2 || P.valid[k] == 0 || P.count[k] <= 0) return;
  const int bx0 = bbox[4 * k], bx1 = bbox[4 * k + 1], by0 = bbox[4 * k + 2], by1 = bbox[4 * k + 3];
  if (bx1 < bx0 || by1 < by0) return;
  const int bw = bx1 - bx0 + 1, bh = by1 - by0 + 1, nc = bw * bh;
  const int s1 = split_seed(P, seg, k, direction, W, H);
  const int s2 = split_seed(P, seg, k, -direction, W, H);
  const int k2 = k + ms + 1;
  __shared__ short d1[kBfsCap], d2[kBfsCap];
  __shared__ uint8_t m[kBfsCap];
  __shared__ int changed;
  const int tid = threadIdx.x;
  if (nc <= kBfsCap) {
    for (int i = tid; i < nc; i += 256) {
      int gx = bx0 + i % bw, gy = by0 + i / bw, g = gy * W + gx;
      m[i] = seg[g] == k;
      d1[i] = (g == s1) ? 1 : -1;
      d2[i] = (g == s2) ? 1 : -1;
    }
    __syncthreads();
    for (int level = 1;; level++) {
      if (tid == 0) changed = 0;
      __syncthreads();
      int ch = 0;
      for (int i = tid; i < nc; i += 256) {
        if (!m[i]) continue;
        bool u1 = d1[i] == -1, u2 = d2[i] == -1;
        if (!u1 && !u2) continue;
        int lx = i % bw, ly = i / bw;
        bool h1 = false, h2 = false;
        if (ly > 0 && m[i - bw]) { h1 |= d1[i - bw] == level; h2 |= d2[i - bw] == level; }
        if (lx > 0 && m[i - 1]) { h1 |= d1[i - 1] == level; h2 |= d2[i - 1] == level; }
        if (ly < bh - 1 && m[i + bw]) { h1 |= d1[i + bw] == level; h2 |= d2[i + bw] == level; }
        if (lx < bw - 1 && m[i + 1]) { h1 |= d1[i + 1] == level; h2 |= d2[i + 1] == level; }
        if (u1 && h1) { d1[i] = (short)(level + 1); ch = 1; }
        if (u2 && h2) { d2[i] = (short)(level + 1); ch = 1; }
      }
      if (ch) changed = 1;
      __syncthreads();
      const int go = changed;
      __syncthreads();  // every thread has read `changed` before it is reset
      if (!go) break;
    }
    for (int i = tid; i < nc; i += 256) {
      if (m[i] && d1[i] > d2[i]) halves[(by0 + i / bw) * W + bx0 + i % bw] = k2;
    }
    return;
  }
  // large label: same levels with the distance maps in global memory
  for (int i = tid; i < nc; i += 256) {
    int g = (by0 + i / bw) * W + bx0 + i % bw;
    if (seg[g] != k) continue;
    d1g[g] = (g == s1) ? 1 : -1;
    d2g[g] = (g == s2) ? 1 : -1;
  }
  __syncthreads();
  for (int level = 1;; level++) {
    if (tid == 0) changed = 0;
    __syncthreads();
    int ch = 0;
    for (int i = tid; i < nc; i += 256) {
      int lx = i % bw, ly = i / bw, g = (by0 + ly) * W + bx0 + lx;
      if (seg[g] != k) continue;
      bool u1 = d1g[g] == -1, u2 = d2g[g] == -1;
      if (!u1 && !u2) continue;
      bool h1 = false, h2 = false;
      int nb[4] = {ly > 0 ? g - W : -1, lx > 0 ? g - 1 : -1, ly < bh - 1 ? g + W : -1,
                   lx < bw - 1 ? g + 1 : -1};
      for (int j = 0; j < 4; j++) {
        int q = nb[j];
        if (q < 0 || seg[q] != k) continue;
        h1 |= d1g[q] == level;
        h2 |= d2g[q] == level;
      }
      if (u1 && h1) { d1g[g] = level + 1; ch = 1; }
      if (u2 && h2) { d2g[g] = level + 1; ch = 1; }
    }
    if (ch) changed = 1;
    __syncthreads();
    const int go = changed;
    __syncthreads();
    if (!go) break;
  }
  for (int i = tid; i < nc; i += 256) {
    int g = (by0 + i / bw) * W + bx0 + i % bw;
    if (seg[g] == k && d1g[g] > d2g[g]) halves[g] = k2;
  }
}

void launch_split_proposal(const int* seg, Params P, const int* bbox, int nbuf,
                           const int* max_spix, int direction, int W, int H, int* d1, int* d2,
                           int* halves, hipStream_t s) {
  int npix = W * H;
  k_copy_i32<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, halves, npix);
  k_bfs_box<<<nbuf, 256, 0, s>>>(seg, P, bbox, nbuf, max_spix, direction, W, H, d1, d2, halves);
}

// ---------------------------------------------------------------------------
// split evaluation (calc_split_stats_step0/1 + update_split_flag, both paths)
// ---------------------------------------------------------------------------
__global__ void k_split_eval(Params P, LabelStats st, LabelStats sst, SmState sm, int nbuf,
                             const int* max_spix, NigConsts nc, float split_alpha,
                             int prop_mode) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int ms = *max_spix;
  if (k >= nbuf || ms > nbuf / 2 || P.valid[k] == 0) return;
  int s = k + (ms + 1);
  if (s >= nbuf) return;
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
  double mk = nig_marginal(nc, (double)count_k, bk);
  double msm = nig_marginal(nc, (double)count_s, bs);
  double mf = nig_marginal(nc, (double)count_f, bf);
  double h = mk + msm - mf + (double)split_alpha;
  if (prop_mode) {
    int is_prop = P.prop[k];
    int nk = sst.ninvalid[k], ns = sst.ninvalid[s];
    int propogate_k = (nk <= ns);
    sm.select[k] = propogate_k || !is_prop;
    double iperc_k = nk / (1.0 * count_k);
    double iperc_s = nk / (1.0 * count_s);
    double iperc = propogate_k ? iperc_k : iperc_s;
    iperc = is_prop ? iperc : 1.0;
    h = h - 2.0 * (1.0 - iperc);
  }
  sm.hasting[k] = h;
  sm.accept[k] = h > -2.0 ? 1 : 0;
}

__global__ void k_split_assign(Params P, LabelStats sst, SmState sm, int nbuf,
                               const int* max_spix, int prop_mode) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int ms = *max_spix;
  if (k >= nbuf || ms > nbuf / 2 || !sm.accept[k]) return;
  int s = ms + 1 + sm.rank[k];
  sm.pairs[2 * k] = s;
  if (!prop_mode) {
    float pc = (float)fmax(P.prior_count[k] / 2.0, 8.0);
    P.prior_count[k] = pc;
    P.prior_count[s] = pc;
  } else {
    int nk = sst.ninvalid[k], ns = sst.ninvalid[k + ms + 1];
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

// split_sp(_p), split_merge.cu:1002-1033, split_merge_prop.cu:2317-2349
__global__ void k_split_apply(int* seg, const int* __restrict__ halves, Params P, SmState sm,
                              int npix, int nbuf, const int* max_spix, int prop_mode) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int ms = *max_spix;
  if (t >= npix || ms > nbuf / 2) return;
  int k = seg[t];
  if (k < 0 || k >= nbuf || !sm.accept[k]) return;
  int k2 = k + (ms + 1);
  int h = halves[t];
  bool take = prop_mode ? (sm.select[k] ? (h == k2) : (h == k)) : (h == k2);
  if (take) {
    int s = sm.pairs[2 * k];
    seg[t] = s;
    P.valid[s] = 1;
  }
}

__global__ void k_add_max(int* max_spix, const int* total, int nbuf, int guard) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (guard && *max_spix > nbuf / 2) return;
    *max_spix += *total;
  }
}

void launch_split_eval_apply(int* seg, const int* halves, Params P, LabelStats st,
                             LabelStats sst, SmState sm, int nbuf, int* max_spix,
                             const NigConsts& nc, float split_alpha, int prop_mode, int npix,
                             int* total, hipStream_t s) {
  k_split_eval<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, st, sst, sm, nbuf, max_spix, nc,
                                                       split_alpha, prop_mode);
  launch_scan(sm.accept, sm.rank, nbuf, total, s);
  k_split_assign<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, sst, sm, nbuf, max_spix, prop_mode);
  k_split_apply<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, halves, P, sm, npix, nbuf, max_spix,
                                                        prop_mode);
  k_add_max<<<1, 1, 0, s>>>(max_spix, total, nbuf, 1);
}

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------
// calc_merge_candidate(_p), split_merge.cu:719-755, split_merge_prop.cu:1808-1845
__global__ void k_merge_cand(const int* __restrict__ seg, const uint8_t* __restrict__ border,
                             const int* __restrict__ prop, int* pairs, int W, int H,
                             int direction, int prop_mode) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= W * H || !border[idx]) return;
  int x = idx % W, y = idx / W;
  int C = seg[idx], Wn = -1;
  if (direction == 1) {
    bool ok = prop_mode ? (y > 1 && y < H - 2) : (y >= 0 && y < H - 1);
    if (ok) Wn = seg[idx + W];
  } else {
    bool ok = prop_mode ? (x > 1 && x < W - 2) : (x > 0 && x < W);
    if (ok) Wn = seg[idx - 1];
  }
  if (Wn >= 0 && C != Wn) {
    if (prop_mode && (prop[C] == prop[Wn])) return;
    atomicMax(pairs + 2 * C + 1, Wn);
  }
}

__global__ void k_merge_step0(Params P, LabelStats st, LabelStats sst, SmState sm, int nbuf,
                              NigConsts nc) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf || P.valid[k] == 0) return;
  int f = sm.pairs[2 * k + 1];
  if (f < 0) return;
  int count_f = P.count[f], count_k = P.count[k];
  int count_fk = count_f + count_k;
  sm.count_f[k] = count_fk;
  for (int c = 0; c < 3; c++) {
    double Qk = fx_to_double(sst.Q[3 * k + c]), Qf = fx_to_double(sst.Q[3 * f + c]);
    double Sk = fx_to_double(st.S[3 * k + c]), Sf = fx_to_double(st.S[3 * f + c]);
    double b = nc.b0 + 0.5 * (Qk - (Sk * Sk) / (double)count_k);
    double bf = nc.b0 + 0.5 * ((Qk + Qf) - ((Sf + Sk) * (Sf + Sk)) / (double)count_fk);
    if (b < 0) b = 0.1;
    if (bf < 0) bf = 0.1;
    sm.b_n[3 * k + c] = b;
    sm.b_n_f[3 * k + c] = bf;
  }
  sm.has_bn[k] = 1;
}

__global__ void k_merge_marg(Params P, SmState sm, int nbuf, NigConsts nc) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf || P.valid[k] == 0) return;
  sm.marg[k] = nig_marginal(nc, (double)P.count[k], sm.b_n + 3 * k);
}

__global__ void k_merge_eval(Params P, SmState sm, int nbuf, NigConsts nc, float merge_alpha) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf || P.valid[k] == 0) return;
  int f = sm.pairs[2 * k + 1];
  if (f < 0 || P.valid[f] == 0) return;
  int count_k = P.count[k], count_fk = sm.count_f[k];
  if (count_k < 1 || count_fk < 1) return;
  double mfk = nig_marginal(nc, (double)count_fk, sm.b_n_f + 3 * k);
  double h = mfk - (nc.log_alpha + sm.marg[k] + sm.marg[f]) + (double)merge_alpha;
  sm.hasting[k] = h;
  if (h > -2.0) atomicMax(sm.winner + f, k);
}

// winner flags + snapshot for the remove step
__global__ void k_merge_flag(Params P, SmState sm, int nbuf) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nbuf) return;
  int w = sm.winner[f];
  if (w >= 0) {
    sm.merge[w] = 1;
    sm.pairs[2 * f] = w;
  }
  sm.valid0[f] = P.valid[f];
  sm.pc0[f] = P.prior_count[f];
  sm.rank[f] = sm.pairs[2 * f + 1];  // target snapshot
}

// remove_sp(_p), split_merge.cu:1035-1080, split_merge_prop.cu:2353-2449
__global__ void k_merge_remove(Params P, SmState sm, int nbuf, int prop_mode) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  const int* tgt0 = sm.rank;
  int s = tgt0[k];
  if (s < 0) return;
  int is_cycle = tgt0[s] == k;
  if (sm.valid0[k] == 0 || sm.valid0[s] == 0) return;
  int mk = sm.merge[k], ms = sm.merge[s];
  if (!prop_mode) {
    if (mk && (!ms || (is_cycle && k < s))) {
      sm.remove[k] = 1;
      P.valid[k] = 0;
      P.prior_count[s] = sm.pc0[k] + sm.pc0[s];
    }
  } else {
    int is_prop_k = P.prop[k];
    int cond_c = is_cycle && is_prop_k;
    if (mk && (!ms || cond_c)) {
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

__global__ void k_merge_apply(int* seg, SmState sm, int npix, int nbuf) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int k = seg[t];
  if (k >= 0 && k < nbuf && sm.remove[k]) seg[t] = sm.pairs[2 * k + 1];
}

void launch_merge(int* seg, const uint8_t* border, Params P, LabelStats st, LabelStats sst,
                  SmState sm, int nbuf, int nsm, const NigConsts& nc, float merge_alpha,
                  int direction, int prop_mode, const float* img, int W, int H, hipStream_t s) {
  int npix = W * H;
  launch_sm_clear(sm, sst, nsm, nbuf, s);
  k_merge_cand<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, border, P.prop, sm.pairs, W, H,
                                                       direction, prop_mode);
  launch_stats(img, seg, nullptr, npix, W, sst, nsm, s);
  k_merge_step0<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, st, sst, sm, nbuf, nc);
  k_merge_marg<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, sm, nbuf, nc);
  k_merge_eval<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, sm, nbuf, nc, merge_alpha);
  k_merge_flag<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, sm, nbuf);
  k_merge_remove<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, sm, nbuf, prop_mode);
  k_merge_apply<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, sm, npix, nbuf);
}

}  // namespace stsp
