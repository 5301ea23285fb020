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
    default: return div_up(W, 32) * div_up(H, 32);  // kProTile
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
// the holes at its start (stop at zero; a count that did not shrink is the
// reference's stall, fill_missing.cu:97-111), marks the border holes (a valid
// 4-neighbour), then fills border holes in four checkerboard phases.  A fill
// never makes a new hole, so the pixels that are holes when a fill starts
// (listed once, in the frame's `keys` buffer, free after the shift) hold
// every hole of its rounds.  One workgroup runs all rounds of a frame, a
// barrier between the stages.
// ---------------------------------------------------------------------------
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

// One workgroup per frame runs the rounds as a frontier.  Every hole that
// has a valid 4-neighbour when a round starts is filled in that round (the
// neighbour stays valid, so its phase finds a candidate), and no other hole
// is; so round r fills exactly its frontier F_r, and F_{r + 1} is the set of
// holes still empty next to F_r.  The hole count at round r is the count at
// the start less the frontiers filled, and a count that does not shrink is an
// empty frontier.  The frontiers alternate between the two halves of the
// frame's `keys` buffer (2 npix ints) after the hole sub-lists are read,
// each frontier as one list per stage of a round; a bit per pixel in the
// `border` buffer marks holes already queued.
constexpr int kFillWaves = 16;
constexpr int kFillThreads = 64 * kFillWaves;

// appends v (when take) to list at the LDS counter *n, one LDS atomic per wave
__device__ __forceinline__ void wave_append(int* list, int* n, bool take, int v) {
  const unsigned long long b = __ballot(take);
  if (!b) return;
  const int lane = threadIdx.x & 63;
  const int first = __ffsll((long long)b) - 1;
  int base = 0;
  if (lane == first) base = atomicAdd(n, __popcll(b));
  base = __shfl(base, first);
  if (take) list[base + __popcll(b & ((1ull << lane) - 1ull))] = v;
}

__global__ void __launch_bounds__(kFillThreads) kp_fill_all(Pro p) {
  const int f = blockIdx.x;
  const ProFrame& F = p.fr[f];
  const int W = F.W, H = F.H, npix = F.npix;
  int* seg = F.seg;
  uint32_t* queued = reinterpret_cast<uint32_t*>(F.border);  // npix bits
  int* const lists = reinterpret_cast<int*>(F.keys);          // 2 npix ints
  const int subcap = hole_sub_cap(npix);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ int ncur, nnext, nholes;
  if (tid == 0) {
    ncur = 0;
    nholes = 0;
  }
  for (int i = tid; i < (npix + 31) / 32; i += kFillThreads) queued[i] = 0u;
  __syncthreads();
  // the holes and the first frontier (holes next to a valid pixel), read from
  // the hole sub-lists into registers before the frontier overwrites them
  {
    constexpr int kSubPerWave = kHoleSub / kFillWaves;
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < kSubPerWave; q++) {
      const int sub = wv + q * kFillWaves;
      const int len = p.hlen[f * kHoleSub + sub];
      cnt += len;
    }
    cnt = lane == 0 ? cnt : 0;
    if (cnt) atomicAdd(&nholes, cnt);
  }
  __syncthreads();
  const int holes0 = nholes;
  // mark the first frontier's holes in the bitmap (the sub-lists still
  // occupy the frontier lists' space), then collect the marked bits
  {
    constexpr int kSubPerWave = kHoleSub / kFillWaves;
    for (int q = 0; q < kSubPerWave; q++) {
      const int sub = wv + q * kFillWaves;
      const int len = p.hlen[f * kHoleSub + sub];
      const int* sl = lists + (size_t)sub * subcap;
      for (int i0 = 0; i0 < len; i0 += 64) {
        const int i = i0 + lane;
        int t = -1;
        bool take = false;
        if (i < len) {
          t = sl[i];
          const int x = t % W, y = t / W;
          take = (y > 0 && seg[t - W] >= 0) || (x > 0 && seg[t - 1] >= 0) ||
                 (y < H - 1 && seg[t + W] >= 0) || (x < W - 1 && seg[t + 1] >= 0);
        }
        if (take) atomicOr(&queued[t >> 5], 1u << (t & 31));
      }
    }
  }
  __syncthreads();
  // every queued bit is a first-frontier hole: collect them from the bitmap
  // into a flat list in the second half of the buffer (the sub-lists are no
  // longer needed)
  int* const flat = lists + npix;
  {
    const int nwords = (npix + 31) / 32;
    for (int w0 = 0; w0 < nwords; w0 += kFillThreads) {
      const int w = w0 + tid;
      uint32_t bits = w < nwords ? queued[w] : 0u;
      const int c = __popc(bits);
      int x = c;  // inclusive scan of the counts over the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
      }
      const int total = __shfl(x, 63);
      int base = 0;
      if (lane == 0 && total) base = atomicAdd(&ncur, total);
      base = __shfl(base, 0);
      int pos = base + x - c;
      while (bits) {
        const int bt = __ffs(bits) - 1;
        bits &= bits - 1;
        flat[pos++] = w * 32 + bt;
      }
    }
  }
  // Each round runs in three stages: the phases (0,0); (0,1) with (1,0),
  // whose cells are never 4-neighbours of each other, so neither reads the
  // other's labels; (1,1).  A frontier is kept as one list per stage: a
  // buffer half holds the stage lists at offsets 0, q0 and q0 + q1, where
  // q0, q1 and npix - q0 - q1 are the frame's cell counts of each stage's
  // phases, so no list can overflow.  A cell filled in a stage appends its
  // empty, unqueued 4-neighbours to the next frontier's list of their stage
  // (a hole of this round's frontier is queued already; a 4-neighbour is
  // never filled in the same stage): neighbours of (0,0) and (1,1) cells are
  // all of stage 1, those of a stage-1 cell of stage 0 or 2 by direction.
  // The counters alternate by round parity: round r reads fcnt[r & 1][..]
  // and appends to the other set, which round r - 1 read and cleared after
  // its first stage.
  const int hx = (W + 1) >> 1, lx = W >> 1, hy = (H + 1) >> 1, ly = H >> 1;
  const int q0 = hx * hy, q1 = hx * ly + lx * hy;
  const int qoff[3] = {0, q0, q0 + q1};
  __shared__ int fcnt[2][3];
  if (tid < 6) fcnt[tid / 3][tid % 3] = 0;
  __syncthreads();
  // the first frontier into the stage lists of the first half
  for (int i0 = 0; i0 < ncur; i0 += kFillThreads) {
    const int i = i0 + tid;
    const int idx = i < ncur ? flat[i] : 0;
    const int y = idx / W, x = idx - y * W;
    const int par = ((x & 1) << 1) | (y & 1);
    const int st = par == 0 ? 0 : par == 3 ? 2 : 1;
#pragma unroll
    for (int t = 0; t < 3; t++) wave_append(lists + qoff[t], &fcnt[0][t], i < ncur && st == t, idx);
  }
  __syncthreads();
  int holes = holes0, rounds = 0, stalled = 0;
  int* cur = lists;
  int* nxt = lists + npix;
  for (int r = 0; r < kFillMax; r++) {
    // this round's counts, read before the first stage's barrier (the
    // counters are cleared after it)
    const int nr[3] = {fcnt[r & 1][0], fcnt[r & 1][1], fcnt[r & 1][2]};
    const int n = nr[0] + nr[1] + nr[2];  // block-uniform
    rounds = r + 1;
    if (holes == 0) break;
    if (n == 0) {  // the count would not shrink
      stalled = 1;
      break;
    }
    int* const cnt_next = fcnt[(r + 1) & 1];
    for (int stage = 0; stage < 3; stage++) {
      const int ns = nr[stage];
      const int* const list = cur + qoff[stage];
      for (int i0 = 0; i0 < ns; i0 += kFillThreads) {
        const int i = i0 + tid;
        const bool mine = i < ns;
        const int idx = mine ? list[i] : 0;
        const int y = idx / W, x = idx - y * W;
        // N, S, W, E
        const int nbv[4] = {mine && y > 0 ? seg[idx - W] : -1, mine && y < H - 1 ? seg[idx + W] : -1,
                            mine && x > 0 ? seg[idx - 1] : -1, mine && x < W - 1 ? seg[idx + 1] : -1};
        // empty 4-neighbours not yet queued join the next frontier: their
        // queue claims go out first (they depend only on the neighbours'
        // labels), then the candidates' centers, all in flight together
        const int nbi[4] = {idx - W, idx + W, idx - 1, idx + 1};
        const bool ok[4] = {mine && y > 0, mine && y < H - 1, mine && x > 0, mine && x < W - 1};
        uint32_t prev[4];
#pragma unroll
        for (int d = 0; d < 4; d++)
          prev[d] = (ok[d] && nbv[d] == -1) ? atomicOr(&queued[nbi[d] >> 5], 1u << (nbi[d] & 31))
                                            : 0xffffffffu;
        double cx[4], cy[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int L = nbv[j] >= 0 ? nbv[j] : 0;
          cx[j] = F.centers[2 * L];
          cy[j] = F.centers[2 * L + 1];
        }
        if (mine) {
          float best = -999999.0f;
          int lab = -1;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int L = nbv[j];
            if (L < 0) continue;
            bool dup = false;
            for (int u = 0; u < j; u++) dup |= nbv[u] == L;
            if (dup) continue;
            // isotropic_space, fill_missing.cu:140-153
            float dx = (float)(((double)(1.0f * (float)x) - cx[j]) / (double)(1.0f * (float)W));
            float dy = (float)(((double)(1.0f * (float)y) - cy[j]) / (double)(1.0f * (float)H));
            float sim = -dx * dx - dy * dy;
            if (sim > best) {
              best = sim;
              lab = L;
            }
          }
          seg[idx] = lab;
        }
        // the appends, in the list of each neighbour's stage: N and S
        // neighbours flip the row parity, W and E the column parity
        const bool ycell = (y & 1) != 0;  // stage 1: a (0,1) cell; else (1,0)
#pragma unroll
        for (int d = 0; d < 4; d++) {
          const bool take = !((prev[d] >> (nbi[d] & 31)) & 1u);
          if (stage != 1) {  // block-uniform
            wave_append(nxt + qoff[1], &cnt_next[1], take, nbi[d]);
          } else {
            // from (0,1): N/S reach (0,0), W/E reach (1,1); from (1,0) the
            // other way round
            const bool to0 = (d < 2) == ycell;
            wave_append(nxt + qoff[0], &cnt_next[0], take && to0, nbi[d]);
            wave_append(nxt + qoff[2], &cnt_next[2], take && !to0, nbi[d]);
          }
        }
      }
      __syncthreads();  // the next stage reads this one's labels
      // every thread read this round's counts before the first barrier, and
      // the next round appends to them only after two more barriers
      if (stage == 0 && tid < 3) fcnt[r & 1][tid] = 0;
    }
    holes -= n;
    int* t = cur;
    cur = nxt;
    nxt = t;
  }
  if (tid == 0) {
    p.fillres[2 * f] = rounds;
    p.fillres[2 * f + 1] = stalled;
  }
}

void pro_fill_all(const Pro& p, hipStream_t s) {
  kp_fill_all<<<p.nall, kFillThreads, 0, s>>>(p);
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
