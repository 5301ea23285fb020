// Label propagation between frames (reference shift_and_fill.cu, sp_pooling.cu,
// shift_labels.cu, fill_missing.cu, split_disconnected.cu, relabel.cu,
// sparams_io.cu).  All collisions and ties are resolved by order-independent
// integer keys (64-bit atomicMin/atomicMax), so the result does not depend on
// thread scheduling.
#include "kernels.h"

namespace stsp {

// ---------------------------------------------------------------------------
// per-label mean flow (run_sp_downsample, sp_pooling.cu:27-71) and shifted
// centres (shift_and_fill.cu:144-150); rule R1 (2^-24 fixed point).
// ---------------------------------------------------------------------------
__global__ void k_flow_clear(int64_t* fsum, int* counts, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  fsum[2 * k] = fsum[2 * k + 1] = 0;
  counts[k] = 0;
}

__global__ void k_flow_accum(const float* __restrict__ flow, const int* __restrict__ spix,
                             int npix, int64_t* fsum, int* counts, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int L = -1;
  int64_t u = 0, v = 0;
  if (t < npix) {
    L = spix[t];
    if (L >= n) L = -1;
    if (L >= 0) {
      u = __double2ll_rn((double)flow[2 * t] * kFxFlow);
      v = __double2ll_rn((double)flow[2 * t + 1] * kFxFlow);
    }
  }
  int run_end;
  bool head = wave_run(L, run_end);
  u = run_sum64(u, run_end);
  v = run_sum64(v, run_end);
  if (head && L >= 0) {
    atomic_add_i64(fsum + 2 * L, u);
    atomic_add_i64(fsum + 2 * L + 1, v);
    atomicAdd(counts + L, run_end - (int)(threadIdx.x & 63));
  }
}

__global__ void k_flow_norm(const int64_t* fsum, const int* counts, const double* mu_shape,
                            float* flow_ds, double* centers, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  for (int c = 0; c < 2; c++) {
    double sum = (double)fsum[2 * k + c] * (1.0 / kFxFlow);
    float v = counts[k] > 0 ? (float)(sum / (1. * counts[k])) : 0.0f;
    flow_ds[2 * k + c] = v;
    centers[2 * k + c] = mu_shape[2 * k + c] + (double)v;
  }
}

void launch_flow_pool(const float* flow, const int* spix, int npix, int64_t* fsum, int* counts,
                      const double* mu_shape, float* flow_ds, double* centers, int n,
                      hipStream_t s) {
  k_flow_clear<<<div_up(n, kBlock), kBlock, 0, s>>>(fsum, counts, n);
  k_flow_accum<<<div_up(npix, kBlock), kBlock, 0, s>>>(flow, spix, npix, fsum, counts, n);
  k_flow_norm<<<div_up(n, kBlock), kBlock, 0, s>>>(fsum, counts, mu_shape, flow_ds, centers, n);
}

// ---------------------------------------------------------------------------
// label shift with collision resolution (shift_labels.cu:32-164), rule R5:
// the destination keeps the lexicographic minimum of (source size, label).
// ---------------------------------------------------------------------------
__global__ void k_u64_fill(unsigned long long* p, unsigned long long v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_shift(const int* __restrict__ spix, const float* __restrict__ flow_ds,
                        const int* __restrict__ sizes, unsigned long long* keys, int W, int H) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= W * H) return;
  int L = spix[t];
  if (L < 0) return;
  int h = t / W, w = t % W;
  int fw = (int)roundf(flow_ds[2 * L]);
  int fh = (int)roundf(flow_ds[2 * L + 1]);
  int wd = w + fw, hd = h + fh;
  if (wd < 0 || wd >= W || hd < 0 || hd >= H) return;
  unsigned long long key = ((unsigned long long)(unsigned)sizes[L] << 32) | (unsigned)L;
  atomicMin(keys + hd * W + wd, key);
}

__global__ void k_shift_decode(const unsigned long long* keys, int* shifted, int npix) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  unsigned long long k = keys[t];
  shifted[t] = (k == ~0ull) ? -1 : (int)(unsigned)(k & 0xffffffffull);
}

void launch_shift(const int* spix, const float* flow_ds, const int* sizes,
                  unsigned long long* keys, int* shifted, int W, int H, hipStream_t s) {
  int npix = W * H;
  k_u64_fill<<<div_up(npix, kBlock), kBlock, 0, s>>>(keys, ~0ull, npix);
  k_shift<<<div_up(npix, kBlock), kBlock, 0, s>>>(spix, flow_ds, sizes, keys, W, H);
  k_shift_decode<<<div_up(npix, kBlock), kBlock, 0, s>>>(keys, shifted, npix);
}

// ---------------------------------------------------------------------------
// hole filling (fill_missing.cu:35-130, 288-441, 617-647)
// ---------------------------------------------------------------------------
__global__ void k_fill_border(const int* __restrict__ seg, uint8_t* border, int* num_neg, int W,
                              int H) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  bool neg = false;
  if (idx < W * H) {
    int x = idx % W, y = idx / W;
    int N = y > 0 ? seg[idx - W] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
    int S = y < H - 1 ? seg[idx + W] : -1, E = x < W - 1 ? seg[idx + 1] : -1;
    border[idx] = (N >= 0) || (Wl >= 0) || (S >= 0) || (E >= 0);
    neg = seg[idx] < 0;
  }
  unsigned long long b = __ballot(neg);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(num_neg, __popcll(b));
}

__global__ void k_fill_phase(int* seg, const uint8_t* __restrict__ border,
                             const double* __restrict__ centers, int W, int H, int xm, int ym) {
  int hw = (W - xm + 1) / 2, hh = (H - ym + 1) / 2;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= hw * hh) return;
  int x = 2 * (i % hw) + xm, y = 2 * (i / hw) + ym;
  int idx = y * W + x;
  if (!border[idx] || seg[idx] != -1) return;
  int N = y > 0 ? seg[idx - W] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
  int S = y < H - 1 ? seg[idx + W] : -1, E = x < W - 1 ? seg[idx + 1] : -1;
  const int cand[4] = {N, S, Wl, E};
  float best = -999999.0f;
  int lab = -1;
  for (int j = 0; j < 4; j++) {
    int L = cand[j];
    if (L < 0) continue;
    bool dup = false;
    for (int q = 0; q < j; q++) dup |= cand[q] == L;
    if (dup) continue;
    // isotropic_space, fill_missing.cu:140-153
    float dx = (float)(((double)(1.0f * (float)x) - centers[2 * L]) / (double)(1.0f * (float)W));
    float dy = (float)(((double)(1.0f * (float)y) - centers[2 * L + 1]) / (double)(1.0f * (float)H));
    float sim = -dx * dx - dy * dy;
    if (sim > best) {
      best = sim;
      lab = L;
    }
  }
  seg[idx] = lab;
}

void launch_fill_count(const int* seg, uint8_t* border, int* num_neg, int W, int H,
                       hipStream_t s) {
  k_fill_border<<<div_up((long long)W * H, kBlock), kBlock, 0, s>>>(seg, border, num_neg, W, H);
}

void launch_fill_phases(int* seg, const uint8_t* border, const double* centers, int W, int H,
                        hipStream_t s) {
  for (int xm = 0; xm < 2; xm++)
    for (int ym = 0; ym < 2; ym++) {
      int hw = (W - xm + 1) / 2, hh = (H - ym + 1) / 2;
      k_fill_phase<<<div_up((long long)hw * hh, kBlock), kBlock, 0, s>>>(seg, border, centers, W,
                                                                         H, xm, ym);
    }
}

// ---------------------------------------------------------------------------
// 8-connected components and removal of all but the largest region of each
// label (split_disconnected.cu:33-185), rule R6.  Union-find with atomicMin
// links every component to its smallest pixel index (Playne & Hawick 2018).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int uf_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int uf_find(const int* parent, int i) {
  int p = uf_load(parent + i);
  while (p != i) {
    i = p;
    p = uf_load(parent + i);
  }
  return i;
}
__device__ __forceinline__ void uf_union(int* parent, int a, int b) {
  while (true) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) {
      int old = atomicMin(parent + b, a);
      if (old == b) return;
      b = old;
    } else {
      int old = atomicMin(parent + a, b);
      if (old == a) return;
      a = old;
    }
  }
}

// Stage 1: one workgroup links the pixels of a kCcTile x kCcTile tile in
// LDS (same rule: every tile component to its smallest pixel index, which is
// also its smallest frame index because both orders are row-major), then
// points each pixel's global parent at its tile root.  Pixels of label < 0
// are left alone: their components never decide anything.
constexpr int kCcTile = 32;

__device__ __forceinline__ int lds_find(int* P, int i) {
  int p = __hip_atomic_load(P + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (p != i) {
    i = p;
    p = __hip_atomic_load(P + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  return i;
}
__device__ __forceinline__ void lds_union(int* P, int a, int b) {
  while (true) {
    a = lds_find(P, a);
    b = lds_find(P, b);
    if (a == b) return;
    if (a < b) {
      int old = atomicMin(P + b, a);
      if (old == b) return;
      b = old;
    } else {
      int old = atomicMin(P + a, b);
      if (old == a) return;
      a = old;
    }
  }
}

__global__ void __launch_bounds__(256) k_cc_tile(const int* __restrict__ seg, int* parent,
                                                 int* rsize, int W, int H) {
  constexpr int T = kCcTile, N = T * T;
  __shared__ int lab[N];
  __shared__ int P[N];
  const int tiles_x = (W + T - 1) / T;
  const int tx0 = (blockIdx.x % tiles_x) * T, ty0 = (blockIdx.x / tiles_x) * T;
  const int tid = threadIdx.x;
  int v[N / 256];
#pragma unroll
  for (int k = 0; k < N / 256; k++) {
    const int i = tid + 256 * k;
    const int gx = tx0 + i % T, gy = ty0 + i / T;
    v[k] = (gx < W && gy < H) ? seg[gy * W + gx] : -1;
  }
#pragma unroll
  for (int k = 0; k < N / 256; k++) {
    const int i = tid + 256 * k;
    lab[i] = v[k];
    P[i] = i;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N / 256; k++) {
    const int i = tid + 256 * k;
    const int lx = i % T, ly = i / T, L = v[k];
    if (L < 0) continue;
    if (lx > 0 && lab[i - 1] == L) lds_union(P, i, i - 1);
    if (ly > 0) {
      if (lab[i - T] == L) lds_union(P, i, i - T);
      if (lx > 0 && lab[i - T - 1] == L) lds_union(P, i, i - T - 1);
      if (lx < T - 1 && lab[i - T + 1] == L) lds_union(P, i, i - T + 1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N / 256; k++) {
    const int i = tid + 256 * k;
    const int gx = tx0 + i % T, gy = ty0 + i / T;
    if (gx >= W || gy >= H) continue;
    const int r = lds_find(P, i);
    parent[gy * W + gx] = (ty0 + r / T) * W + tx0 + r % T;
    rsize[gy * W + gx] = 0;
  }
}

// Stage 2: the neighbour pairs that cross a tile edge (each pixel's W, N, NW
// and NE neighbours; a pair is owned by its later pixel) are linked in the
// global forest, whose chains now run only between tile roots.
__global__ void __launch_bounds__(128) k_cc_seams(const int* __restrict__ seg, int* parent, int W,
                                                  int H) {
  constexpr int T = kCcTile;
  const int tiles_x = (W + T - 1) / T;
  const int tx0 = (blockIdx.x % tiles_x) * T, ty0 = (blockIdx.x / tiles_x) * T;
  const int t = threadIdx.x;
  if (t >= 3 * T) return;
  // t < T: top row; T <= t < 2T: left column; 2T <= t: right column
  const int lx = t < T ? t : (t < 2 * T ? 0 : T - 1);
  const int ly = t < T ? 0 : t % T;
  const int x = tx0 + lx, y = ty0 + ly;
  if (x >= W || y >= H) return;
  const int i = y * W + x, L = seg[i];
  if (L < 0) return;
  const bool top = ly == 0, left = lx == 0, right = lx == T - 1;
  if (t >= T && top) return;  // the top row owns its corners
  if (left && x > 0 && seg[i - 1] == L) uf_union(parent, i, i - 1);
  if (y > 0) {
    if (top && seg[i - W] == L) uf_union(parent, i, i - W);
    if ((top || left) && x > 0 && seg[i - W - 1] == L) uf_union(parent, i, i - W - 1);
    if ((top || right) && x < W - 1 && seg[i - W + 1] == L) uf_union(parent, i, i - W + 1);
  }
}

// Points every pixel at its root and counts region sizes; neighbouring
// pixels of a wave mostly share a root, so the counts are added once per run
// of equal roots.
__global__ void k_cc_compress(int* parent, int* rsize, int npix) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int r = -1;
  if (i < npix) {
    r = uf_find(parent, i);
    parent[i] = r;
  }
  int run_end;
  const bool head = wave_run(r, run_end);
  if (head && r >= 0) atomicAdd(rsize + r, run_end - (int)(threadIdx.x & 63));
}

__global__ void k_cc_best(const int* __restrict__ seg, const int* __restrict__ root,
                          const int* __restrict__ rsize, unsigned long long* best, int npix,
                          int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix || root[i] != i) return;
  int L = seg[i];
  if (L < 0 || L >= n) return;
  unsigned long long key = ((unsigned long long)(unsigned)rsize[i] << 32) |
                           (unsigned long long)(0xffffffffu - (unsigned)i);
  atomicMax(best + L, key);
}

__global__ void k_cc_apply(int* seg, const int* __restrict__ root,
                           const unsigned long long* __restrict__ best, int npix, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  int L = seg[i];
  if (L < 0 || L >= n) return;
  unsigned keep = 0xffffffffu - (unsigned)(best[L] & 0xffffffffull);
  if ((unsigned)root[i] != keep) seg[i] = -1;
}

void launch_invalidate_disconnected(int* seg, int* parent, int* rsize, unsigned long long* best,
                                    int W, int H, int n, hipStream_t s) {
  int npix = W * H;
  const int tiles = div_up(W, kCcTile) * div_up(H, kCcTile);
  k_u64_fill<<<div_up(n, kBlock), kBlock, 0, s>>>(best, 0ull, n);
  k_cc_tile<<<tiles, 256, 0, s>>>(seg, parent, rsize, W, H);
  k_cc_seams<<<tiles, 128, 0, s>>>(seg, parent, W, H);
  k_cc_compress<<<div_up(npix, kBlock), kBlock, 0, s>>>(parent, rsize, npix);
  k_cc_best<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, parent, rsize, best, npix, n);
  k_cc_apply<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, parent, best, npix, n);
}

// ---------------------------------------------------------------------------
// per-label pixel counts, invalid fractions and prior counts
// (shift_and_fill.cu:223-322)
// ---------------------------------------------------------------------------
__global__ void k_count_labels(const int* __restrict__ seg, int npix, int* counts, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int k = -1;
  if (t < npix) {
    k = seg[t];
    if (k >= n) k = -1;
  }
  int run_end;
  bool head = wave_run(k, run_end);
  if (head && k >= 0) atomicAdd(counts + k, run_end - (int)(threadIdx.x & 63));
}

void launch_count_labels(const int* seg, int npix, int* counts, int n, hipStream_t s) {
  launch_fill(counts, 0, n, s);
  k_count_labels<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, npix, counts, n);
}

__global__ void k_prior_update(const int* counts_prev, const int* counts_s, const int* counts_f,
                               float* prior_counts, float* invalid_counts, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int f = counts_f[k], sh = counts_s[k];
  invalid_counts[k] = (float)((double)(f - sh) / ((double)(float)f + 1e-8));
  int c = counts_prev[k];
  if (c > 0) prior_counts[k] = (float)fmin(fmax(1.0 * c, 100.0), 10000.0);
}

void launch_prior_update(const int* counts_prev, const int* counts_s, const int* counts_f,
                         float* prior_counts, float* invalid_counts, int n, hipStream_t s) {
  k_prior_update<<<div_up(n, kBlock), kBlock, 0, s>>>(counts_prev, counts_s, counts_f,
                                                      prior_counts, invalid_counts, n);
}

// ---------------------------------------------------------------------------
// get_vectors_as_params + mark_active (sparams_io.cu:123-140, 352-501;
// init_sparams.cu:195-286)
// ---------------------------------------------------------------------------
__global__ void k_prop_params_init(Params P, InParams in, int nspix, int nbuf, int sp_size) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  if (k < nspix) {
    for (int c = 0; c < 3; c++) {
      P.mu_app[3 * k + c] = in.mu_app[3 * k + c];
      P.prior_mu_app[3 * k + c] = in.prior_mu_app[3 * k + c];
      P.sigma_shape[3 * k + c] = in.sigma_shape[3 * k + c];
      P.prior_sigma_shape[3 * k + c] = in.prior_sigma_shape[3 * k + c];
      P.sample_sigma_shape[3 * k + c] = in.sample_sigma_shape[3 * k + c];
    }
    for (int c = 0; c < 2; c++) {
      P.mu_shape[2 * k + c] = in.mu_shape[2 * k + c];
      P.prior_mu_shape[2 * k + c] = in.prior_mu_shape[2 * k + c];
    }
    P.logdet[k] = (double)in.logdet_sigma_shape[k];
    const double* cv = in.prior_sigma_shape + 3 * k;
    double pr_det = cv[0] * cv[2] - cv[1] * cv[1];
    P.prior_icov[3 * k] = cv[2] / pr_det;
    P.prior_icov[3 * k + 1] = cv[1] / pr_det;
    P.prior_icov[3 * k + 2] = cv[0] / pr_det;
    P.count[k] = in.counts[k];
    P.icount[k] = in.invalid_counts[k];
    P.sm_count[k] = in.sm_counts[k];
    P.prior_count[k] = in.prior_counts[k];
  } else {
    for (int c = 0; c < 3; c++) {
      P.mu_app[3 * k + c] = 0.f;
      P.prior_mu_app[3 * k + c] = 0.f;
      P.sigma_shape[3 * k + c] = 0.0;
      P.sample_sigma_shape[3 * k + c] = 0.0;
      P.prior_icov[3 * k + c] = 0.0;
    }
    P.prior_sigma_shape[3 * k] = 1.0 / sp_size;
    P.prior_sigma_shape[3 * k + 1] = 0;
    P.prior_sigma_shape[3 * k + 2] = 1.0 / sp_size;
    for (int c = 0; c < 2; c++) {
      P.mu_shape[2 * k + c] = 0.0;
      P.prior_mu_shape[2 * k + c] = 0.0;
    }
    P.logdet[k] = 0.0;
    P.count[k] = 0;
    P.icount[k] = 0.f;
    P.sm_count[k] = 0;
    P.prior_count[k] = 0.f;
  }
  P.valid[k] = 0;
  P.prop[k] = 0;
}

__global__ void k_mark_active_px(const int* __restrict__ seg, int npix, Params P, int nbuf) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int L = seg[t];
  if (L >= 0 && L < nbuf) {
    P.valid[L] = 1;
    P.prop[L] = 1;
  }
}

void launch_prop_params_init(Params P, InParams in, int nspix, int nbuf, int sp_size,
                             const int* filled, int npix, hipStream_t s) {
  k_prop_params_init<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, in, nspix, nbuf, sp_size);
  k_mark_active_px<<<div_up(npix, kBlock), kBlock, 0, s>>>(filled, npix, P, nbuf);
}

}  // namespace stsp
