// Label propagation between frames (reference shift_and_fill.cu, sp_pooling.cu,
// shift_labels.cu, fill_missing.cu, split_disconnected.cu, sparams_io.cu).
// All collisions and ties are resolved by order-independent integer keys
// (64-bit atomicMin/atomicMax), so the result does not depend on thread
// scheduling.
//
// Every kernel covers all frames of a batch: block b belongs to the frame f
// with off[kind][f] <= b < off[kind][f + 1], so a stage is one launch however
// many frames there are, and a block never straddles two frames.
#include "kernels.h"

namespace stsp {

__device__ __forceinline__ int pro_frame(const Pro& p, int kind, int b) {
  const int* off = p.off + kind * (kProMax + 1);
  int lo = 0, hi = p.nf - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ int pro_local(const Pro& p, int kind, int f, int b) {
  return b - p.off[kind * (kProMax + 1) + f];
}

// per-pixel kernels: f = frame, t = pixel index (may be >= npix)
#define PRO_PIX                                              \
  const int f = pro_frame(p, kProPix, blockIdx.x);           \
  const ProFrame& F = p.fr[f];                               \
  const int t = pro_local(p, kProPix, f, blockIdx.x) * kBlock + threadIdx.x;
// per-label kernels over the previous labels: k < np
#define PRO_LAB                                              \
  const int f = pro_frame(p, kProLab, blockIdx.x);           \
  const ProFrame& F = p.fr[f];                               \
  const int k = pro_local(p, kProLab, f, blockIdx.x) * kBlock + threadIdx.x;

// adds the block's count of true predicates to *dst with one atomic (all
// threads of the block must call it)
__device__ __forceinline__ void block_count_add(bool pred, int* dst) {
  __shared__ int wc[kBlock / 64];
  const unsigned long long b = __ballot(pred);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int w = 0; w < kBlock / 64; w++) n += wc[w];
    if (n) atomicAdd(dst, n);
  }
}

int pro_blocks(int kind, int W, int H, int np, int nbuf) {
  switch (kind) {
    case kProPix: return div_up((long long)W * H, kBlock);
    case kProLab: return div_up(np, kBlock);
    case kProBuf: return div_up(nbuf, kBlock);
    case kProTile: return div_up(W, 32) * div_up(H, 32);
    case kProHole: {
      const int nb = div_up((long long)W * H, kBlock);
      return nb < kHoleSub ? nb : kHoleSub;
    }
    default: return div_up((long long)((W + 1) / 2) * ((H + 1) / 2), kBlock);  // kProHalf
  }
}

// ---------------------------------------------------------------------------
// per-label mean flow (run_sp_downsample, sp_pooling.cu:27-71) and shifted
// centres (shift_and_fill.cu:144-150); rule R1 (2^-24 fixed point).  The
// clear also zeroes the label counts, the best-region keys of the
// connectivity pass and the hole count, and sets max_spix (prop.cu:736).
// ---------------------------------------------------------------------------
__global__ void kp_clear(Pro p) {
  PRO_LAB
  if (k < kNegSub) p.negs[f * kNegSub + k] = 0;
  if (k == 0) *F.max_spix = F.np - 1;
  if (k >= F.np) return;
  F.fsum[2 * k] = F.fsum[2 * k + 1] = 0;
  F.counts_prev[k] = 0;
  F.counts_s[k] = 0;
  F.counts_f[k] = 0;
  F.best[k] = 0ull;
}

__global__ void kp_flow_accum(Pro p) {
  PRO_PIX
  int L = -1;
  int64_t u = 0, v = 0;
  if (t < F.npix) {
    L = F.spix_prev[t];
    if (L >= F.np) L = -1;
    if (L >= 0) {
      u = __double2ll_rn((double)F.flow[2 * t] * kFxFlow);
      v = __double2ll_rn((double)F.flow[2 * t + 1] * kFxFlow);
    }
  }
  int run_end;
  bool head = wave_run(L, run_end);
  u = run_sum64(u, run_end);
  v = run_sum64(v, run_end);
  if (head && L >= 0) {
    atomic_add_i64(F.fsum + 2 * L, u);
    atomic_add_i64(F.fsum + 2 * L + 1, v);
    atomicAdd(F.counts_prev + L, run_end - (int)(threadIdx.x & 63));
  }
}

__global__ void kp_flow_norm(Pro p) {
  PRO_LAB
  if (k >= F.np) return;
  for (int c = 0; c < 2; c++) {
    double sum = (double)F.fsum[2 * k + c] * (1.0 / kFxFlow);
    float v = F.counts_prev[k] > 0 ? (float)(sum / (1. * F.counts_prev[k])) : 0.0f;
    F.flow_ds[2 * k + c] = v;
    F.centers[2 * k + c] = F.mu_shape_prev[2 * k + c] + (double)v;
  }
}

// ---------------------------------------------------------------------------
// label shift with collision resolution (shift_labels.cu:32-164), rule R5:
// the destination keeps the lexicographic minimum of (source size, label).
// ---------------------------------------------------------------------------
__global__ void kp_keys_fill(Pro p) {
  PRO_PIX
  if (t < F.npix) F.keys[t] = ~0ull;
}

__global__ void kp_shift(Pro p) {
  PRO_PIX
  if (t >= F.npix) return;
  const int W = F.W, H = F.H;
  int L = F.spix_prev[t];
  if (L < 0 || L >= F.np) return;
  int h = t / W, w = t % W;
  int fw = (int)roundf(F.flow_ds[2 * L]);
  int fh = (int)roundf(F.flow_ds[2 * L + 1]);
  int wd = w + fw, hd = h + fh;
  if (wd < 0 || wd >= W || hd < 0 || hd >= H) return;
  unsigned long long key = ((unsigned long long)(unsigned)F.counts_prev[L] << 32) | (unsigned)L;
  atomicMin(F.keys + hd * W + wd, key);
}

// decodes the shifted map into both the kept copy and the working labels
__global__ void kp_shift_decode(Pro p) {
  PRO_PIX
  if (t >= F.npix) return;
  unsigned long long k = F.keys[t];
  const int v = (k == ~0ull) ? -1 : (int)(unsigned)(k & 0xffffffffull);
  F.shifted[t] = v;
  F.seg[t] = v;
}

// per-label pixel counts (shift_and_fill.cu:223-322): of the shifted map
// (which == 0) or of the filled labels (which == 1)
__global__ void kp_count_labels(Pro p, int which) {
  PRO_PIX
  const int* seg = which == 0 ? F.shifted : F.seg;
  int* counts = which == 0 ? F.counts_s : F.counts_f;
  int k = -1;
  if (t < F.npix) {
    k = seg[t];
    if (k >= F.np) k = -1;
  }
  int run_end;
  bool head = wave_run(k, run_end);
  if (head && k >= 0) atomicAdd(counts + k, run_end - (int)(threadIdx.x & 63));
}

// holes of the shifted map (the adaptive iteration count, pyapi.cu:139-149)
__global__ void kp_count_holes(Pro p) {
  PRO_PIX
  // consecutive blocks of a frame take consecutive partial counters
  block_count_add(t < F.npix && F.shifted[t] < 0,
                  p.negs + f * kNegSub + ((t / kBlock) & (kNegSub - 1)));
}

void pro_shift(const Pro& p, const int* total, hipStream_t s) {
  kp_clear<<<total[kProLab], kBlock, 0, s>>>(p);
  kp_flow_accum<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_flow_norm<<<total[kProLab], kBlock, 0, s>>>(p);
  kp_keys_fill<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_shift<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_shift_decode<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_count_labels<<<total[kProPix], kBlock, 0, s>>>(p, 0);
  kp_count_holes<<<total[kProPix], kBlock, 0, s>>>(p);
}

// ---------------------------------------------------------------------------
// hole filling (fill_missing.cu:35-130, 288-441, 617-647).  Round r counts
// the holes and marks the border (a valid 4-neighbour) at its start, then
// fills border holes in four checkerboard phases.  A fill never makes a new
// hole, so the pixels that are holes when a fill starts (listed once, in the
// frame's `keys` buffer, free after the shift) hold every hole of its
// rounds: the rounds visit the list, not the frame.  A frame whose count was
// zero in the round before is finished: its blocks return at once.
// ---------------------------------------------------------------------------
// adds the block's sum of v to *dst with one atomic (all threads call it)
__device__ __forceinline__ void block_sum_add(int v, int* dst) {
  __shared__ int ws[kBlock / 64];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int n = 0;
    for (int w = 0; w < kBlock / 64; w++) n += ws[w];
    if (n) atomicAdd(dst, n);
  }
}

__global__ void kp_hole_clear(Pro p) {
  for (int i = threadIdx.x; i < p.nall * kHoleSub; i += blockDim.x) p.hlen[i] = 0;
}

// entries of one hole sub-list of a frame of npix pixels: the pixel blocks
// that append to it, 256 pixels each
__device__ __forceinline__ int hole_sub_cap(int npix) {
  return ((npix + kBlock - 1) / kBlock + kHoleSub - 1) / kHoleSub * kBlock;
}

// the holes of every frame (labels < 0), appended with one atomic per wave
__global__ void kp_hole_list(Pro p) {
  PRO_PIX
  const bool hole = t < F.npix && F.seg[t] < 0;
  const unsigned long long b = __ballot(hole);
  if (!b) return;
  const int sub = pro_local(p, kProPix, f, blockIdx.x) % kHoleSub;
  const int lane = threadIdx.x & 63;
  const int first = __ffsll((long long)b) - 1;
  int base = 0;
  if (lane == first) base = atomicAdd(p.hlen + f * kHoleSub + sub, __popcll(b));
  base = __shfl(base, first);
  if (hole) {
    int* list = reinterpret_cast<int*>(F.keys) + (size_t)sub * hole_sub_cap(F.npix);
    list[base + __popcll(b & ((1ull << lane) - 1ull))] = t;
  }
}

void pro_hole_list(const Pro& p, const int* total, hipStream_t s) {
  kp_hole_clear<<<1, 1024, 0, s>>>(p);
  kp_hole_list<<<total[kProPix], kBlock, 0, s>>>(p);
}

// The filling rounds are launched over the frames that still have holes
// (p.fmap lists them; the counts keep the table's frame index).
__global__ void kp_fill_zero(Pro p, int base, int rounds) {
  const int n = rounds * p.nall;
  for (int i = threadIdx.x; i < n; i += blockDim.x) p.fillcnt[base * p.nall + i] = 0;
}

// the blocks of frame fl of a hole-list launch: block hb of nhb visits the
// sub-lists hb, hb + nhb, ... (entries i of sub-list sub: HOLE_AT(sub, i))
#define PRO_HOLE                                                           \
  const int fl = pro_frame(p, kProHole, blockIdx.x);                       \
  const int f = p.fmap ? p.fmap[fl] : fl;                                  \
  const ProFrame& F = p.fr[f];                                             \
  const int* off_h = p.off + kProHole * (kProMax + 1);                     \
  const int hb = blockIdx.x - off_h[fl], nhb = off_h[fl + 1] - off_h[fl];  \
  const int* list = reinterpret_cast<const int*>(F.keys);                  \
  const int subcap = hole_sub_cap(F.npix);                                 \
  const int* hl = p.hlen + f * kHoleSub;

__global__ void kp_fill_border(Pro p, int r) {
  PRO_HOLE
  if (r > 0 && p.fillcnt[(r - 1) * p.nall + f] == 0) return;  // block-uniform
  const int W = F.W, H = F.H;
  const int* seg = F.seg;
  int neg = 0;
  for (int sub = hb; sub < kHoleSub; sub += nhb) {
    for (int i = threadIdx.x; i < hl[sub]; i += kBlock) {
      const int t = list[(size_t)sub * subcap + i];
      if (seg[t] >= 0) continue;  // filled in an earlier round
      const int x = t % W, y = t / W;
      const int N = y > 0 ? seg[t - W] : -1, Wl = x > 0 ? seg[t - 1] : -1;
      const int S = y < H - 1 ? seg[t + W] : -1, E = x < W - 1 ? seg[t + 1] : -1;
      F.border[t] = (N >= 0) || (Wl >= 0) || (S >= 0) || (E >= 0);
      neg++;
    }
  }
  block_sum_add(neg, p.fillcnt + r * p.nall + f);
}

__global__ void kp_fill_phase(Pro p, int r, int xm, int ym) {
  PRO_HOLE
  if (p.fillcnt[r * p.nall + f] == 0) return;
  const int W = F.W, H = F.H;
  int* seg = F.seg;
  for (int sub = hb; sub < kHoleSub; sub += nhb) {
    for (int i = threadIdx.x; i < hl[sub]; i += kBlock) {
      const int idx = list[(size_t)sub * subcap + i];
      const int x = idx % W, y = idx / W;
      if ((x & 1) != xm || (y & 1) != ym) continue;
      if (seg[idx] != -1 || !F.border[idx]) continue;
      const int N = y > 0 ? seg[idx - W] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
      const int S = y < H - 1 ? seg[idx + W] : -1, E = x < W - 1 ? seg[idx + 1] : -1;
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
        float dx = (float)(((double)(1.0f * (float)x) - F.centers[2 * L]) / (double)(1.0f * (float)W));
        float dy =
            (float)(((double)(1.0f * (float)y) - F.centers[2 * L + 1]) / (double)(1.0f * (float)H));
        float sim = -dx * dx - dy * dy;
        if (sim > best) {
          best = sim;
          lab = L;
        }
      }
      seg[idx] = lab;
    }
  }
}
#undef PRO_HOLE

void pro_fill_rounds(const Pro& p, const int* total, int base, int rounds, hipStream_t s) {
  kp_fill_zero<<<1, 256, 0, s>>>(p, base, rounds);
  for (int r = base; r < base + rounds; r++) {
    kp_fill_border<<<total[kProHole], kBlock, 0, s>>>(p, r);
    for (int xm = 0; xm < 2; xm++)
      for (int ym = 0; ym < 2; ym++)
        kp_fill_phase<<<total[kProHole], kBlock, 0, s>>>(p, r, xm, ym);
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

__global__ void __launch_bounds__(256) kp_cc_tile(Pro p) {
  constexpr int T = kCcTile, N = T * T;
  __shared__ int lab[N];
  __shared__ int P[N];
  const int f = pro_frame(p, kProTile, blockIdx.x);
  const ProFrame& F = p.fr[f];
  const int W = F.W, H = F.H;
  const int* __restrict__ seg = F.seg;
  const int tile = pro_local(p, kProTile, f, blockIdx.x);
  const int tiles_x = (W + T - 1) / T;
  const int tx0 = (tile % tiles_x) * T, ty0 = (tile / tiles_x) * T;
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
    F.parent[gy * W + gx] = (ty0 + r / T) * W + tx0 + r % T;
    F.rsize[gy * W + gx] = 0;
  }
}

// Stage 2: the neighbour pairs that cross a tile edge (each pixel's W, N, NW
// and NE neighbours; a pair is owned by its later pixel) are linked in the
// global forest, whose chains now run only between tile roots.
__global__ void __launch_bounds__(128) kp_cc_seams(Pro p) {
  constexpr int T = kCcTile;
  const int f = pro_frame(p, kProTile, blockIdx.x);
  const ProFrame& F = p.fr[f];
  const int W = F.W, H = F.H;
  const int* __restrict__ seg = F.seg;
  const int tile = pro_local(p, kProTile, f, blockIdx.x);
  const int tiles_x = (W + T - 1) / T;
  const int tx0 = (tile % tiles_x) * T, ty0 = (tile / tiles_x) * T;
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
  int* parent = F.parent;
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
__global__ void kp_cc_compress(Pro p) {
  PRO_PIX
  int r = -1;
  if (t < F.npix) {
    r = uf_find(F.parent, t);
    F.parent[t] = r;
  }
  int run_end;
  const bool head = wave_run(r, run_end);
  if (head && r >= 0) atomicAdd(F.rsize + r, run_end - (int)(threadIdx.x & 63));
}

__global__ void kp_cc_best(Pro p) {
  PRO_PIX
  if (t >= F.npix || F.parent[t] != t) return;
  int L = F.seg[t];
  if (L < 0 || L >= F.np) return;
  unsigned long long key = ((unsigned long long)(unsigned)F.rsize[t] << 32) |
                           (unsigned long long)(0xffffffffu - (unsigned)t);
  atomicMax(F.best + L, key);
}

__global__ void kp_cc_apply(Pro p) {
  PRO_PIX
  if (t >= F.npix) return;
  int L = F.seg[t];
  if (L < 0 || L >= F.np) return;
  unsigned keep = 0xffffffffu - (unsigned)(F.best[L] & 0xffffffffull);
  if ((unsigned)F.parent[t] != keep) F.seg[t] = -1;
}

void pro_disconnected(const Pro& p, const int* total, hipStream_t s) {
  kp_cc_tile<<<total[kProTile], 256, 0, s>>>(p);
  kp_cc_seams<<<total[kProTile], 128, 0, s>>>(p);
  kp_cc_compress<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_cc_best<<<total[kProPix], kBlock, 0, s>>>(p);
  kp_cc_apply<<<total[kProPix], kBlock, 0, s>>>(p);
}

// ---------------------------------------------------------------------------
// invalid fractions and prior counts (shift_and_fill.cu:223-322), then
// get_vectors_as_params + mark_active (sparams_io.cu:123-140, 352-501;
// init_sparams.cu:195-286)
// ---------------------------------------------------------------------------
__global__ void kp_prior_update(Pro p) {
  PRO_LAB
  if (k >= F.np) return;
  int fc = F.counts_f[k], sh = F.counts_s[k];
  F.invalid_counts[k] = (float)((double)(fc - sh) / ((double)(float)fc + 1e-8));
  int c = F.counts_prev[k];
  if (c > 0) F.prior_counts[k] = (float)fmin(fmax(1.0 * c, 100.0), 10000.0);
}

__global__ void kp_params_init(Pro p) {
  const int f = pro_frame(p, kProBuf, blockIdx.x);
  const ProFrame& F = p.fr[f];
  const int k = pro_local(p, kProBuf, f, blockIdx.x) * kBlock + threadIdx.x;
  if (k >= F.nbuf) return;
  const Params& P = F.P;
  const InParams& in = F.in;
  if (k < F.np) {
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
    P.prior_sigma_shape[3 * k] = 1.0 / F.sp_size;
    P.prior_sigma_shape[3 * k + 1] = 0;
    P.prior_sigma_shape[3 * k + 2] = 1.0 / F.sp_size;
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

__global__ void kp_mark_active(Pro p) {
  PRO_PIX
  if (t >= F.npix) return;
  int L = F.seg[t];
  if (L >= 0 && L < F.nbuf) {
    F.P.valid[L] = 1;
    F.P.prop[L] = 1;
  }
}

void pro_finish(const Pro& p, const int* total, hipStream_t s) {
  kp_count_labels<<<total[kProPix], kBlock, 0, s>>>(p, 1);
  kp_prior_update<<<total[kProLab], kBlock, 0, s>>>(p);
  kp_params_init<<<total[kProBuf], kBlock, 0, s>>>(p);
  kp_mark_active<<<total[kProPix], kBlock, 0, s>>>(p);
}

}  // namespace stsp
