// Split step of the batched refinement loop (split_merge_orig.cu:202-420,
// split_merge_prop.cu:184-395): work lists by box size, the two-seed
// bit-parallel BFS proposals, FP64 Hastings ratios, prefix-sum ids.
#include "kb_dev.h"

namespace stsp {

// Labels that take part in the split (valid, non-empty, split not skipped),
// sorted by the size of their bounding box: boxes up to wave_box x wave_box
// (at most 128 x 128, the wave kernel's register rows) go to the wave list
// work[0..); boxes up to 448 x 384 or 384 x 448 that also fit the frame's
// search window to the block list work[work_cap..) (searched whole by the
// 8-wave kernel); the others to the big-block list work[2 work_cap..).  The
// block list holds its larger boxes (area at least kMidHeavy pixels) from
// its start and the others from its end, so the 8-wave kernel takes the
// longest searches first and the short ones fill in behind them.

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
  } else if (((bh <= kMidRows && bw <= kMidCols) || (bw <= kMidRows && bh <= kMidCols)) &&
             bh <= F.win_h && bw <= F.win_w) {
    if (bw * bh >= kMidHeavy) {
      const int pos = atomicAdd(b.work_n + 2, 1);
      if (pos < b.work_cap) b.work[b.work_cap + pos] = (f << 24) | k;
    } else {
      const int pos = atomicAdd(b.work_n + kWorkMidLight, 1);
      if (pos < b.work_cap) b.work[2 * b.work_cap - 1 - pos] = (f << 24) | k;
    }
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
//
// The band searches (bfs_bands) grow the two regions R1 = {d1 <= d2} and
// R2 = {d2 < d1} instead, each only into pixels neither holds yet (a pixel
// both reach at one level joins R1):
//   N1 = grow(R1) & M & ~(R1 | R2),  N2 = grow(R2) & M & ~(R1 | R2) & ~N1.
// Both sets hold the predecessor of each of their pixels on a shortest path
// from their seed (if d1(p) <= d2(p) then d1(q) = d1(p) - 1 <= d2(q) for
// that q; likewise with < for R2), so level L adds exactly the pixels with
// min(d1, d2) = L.  When the seeds share a 4-connected piece, the new label
// takes R2 (d2 < d1); otherwise seed 2 reaches no pixel of seed 1's piece
// and the new label takes all of R1.  The seeds share a piece exactly when
// R1 and R2 hold 4-adjacent pixels.  Each region grows only in the cells its
// front passes, about half of those either front passes.
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
// round was already a fixed point.  The 512-thread kernel of the block list
// holds 8 bands (448 rows) of up to 6 words (384 columns), one band per
// wave; the 1024-thread kernel of larger boxes 16 bands (896 rows) of 8 words
// (512 columns), one band per wave (four waves per SIMD, so at most 128
// VGPRs: the mask lives in LDS).
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
constexpr int kBigRows = 16 * kOwn, kBigCols = 8 * 64;  // 1024 threads, 16 bands
static_assert(kSplitWinRows == kBigRows && kSplitWinCols == kBigCols,
              "the widest search window is the 1024-thread band layout's");
static_assert(kMidRows == 8 * kOwn && kMidCols == 6 * 64,
              "boxes of the block list fit 8 bands of 6 words");
constexpr int kBlkWords = 2048;
constexpr int kBigWords = 18 * 512;  // whole-box words of the 512-thread kernel (147 KB of LDS)
enum { kEscLeft = 1, kEscRight = 2, kEscTop = 4, kEscBottom = 8 };

template <int NBANDS, int WDMAX, int NW>
struct BandEdge {
  uint64_t v[2][NBANDS][2][kHalo][2][WDMAX];  // [buf][band][top|bottom][row][seed][word]
  uint32_t g[2][NW];                          // [buf][wave]: own rows grew
  uint32_t esc;                               // seed 1's component leaves the window
  uint32_t same;                              // the seeds lie in one piece
};
typedef BandEdge<16, 8, 16> BigEdge;

// Searches the window (bx0, by0, bw, bh) of label k.  esc_sides marks the
// window edges beyond which the label's box continues.  Returns true
// (block-uniform) when seed 1's component reaches such an edge with a pixel of
// label k just outside; the halves map is then left unchanged.
// With MLDS the label mask lives in LDS (mlds, [band][word][lane], 8 * 64 *
// NW * RB * WD bytes, lanes on consecutive words) instead of registers.
// With TR the box is searched transposed: the bands run over its columns
// (a lane holds one column) and the words over its rows (bit r of word w is
// row 64 w + r); 4-connectivity is the same either way.  Transposed
// searches take no window (esc_sides == 0).
template <int NW, int RB, int WD, bool MLDS, class Edge, bool TR = false>
__device__ __forceinline__ bool bfs_bands(const int* __restrict__ seg, int W, int k, int bx0,
                                          int by0, int bw, int bh, int s1, int s2, int k2,
                                          int* __restrict__ halves, Edge& E, int esc_sides,
                                          uint64_t* __restrict__ mlds) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nrow = TR ? bw : bh;  // rows of the band structure
  uint64_t M[MLDS ? 1 : RB][MLDS ? 1 : WD], V1[RB][WD], V2[RB][WD];
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
      V1[c][w] = V2[c][w] = 0;
      const int col = 64 * w + lane;
      if (TR && row0 < nrow) {
        // lane = box column x, bit r = box row 64 w + r: each load reads 64
        // neighbouring columns of one row
        const int x = row0 + lane;
        const bool xin = x >= 0 && x < bw;
        for (int r0 = 0; r0 < 64; r0 += 16) {
          int v[16];
#pragma unroll
          for (int j = 0; j < 16; j++) {
            const int y = 64 * w + r0 + j;
            v[j] = (xin && y < bh) ? seg[(by0 + y) * W + bx0 + x] : -1;
          }
#pragma unroll
          for (int j = 0; j < 16; j++) m |= (uint64_t)(v[j] == k) << (r0 + j);
        }
      } else if (!TR && row0 < bh) {
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
    // seed cells in band coordinates (row y, column x)
    const int iy1 = s1 / W - by0, ix1 = s1 % W - bx0;
    const int iy2 = s2 >= 0 ? s2 / W - by0 : -1000, ix2 = s2 >= 0 ? s2 % W - bx0 : -1000;
    const int y1 = TR ? ix1 : iy1, x1 = TR ? iy1 : ix1;
    const int y2 = s2 >= 0 ? (TR ? ix2 : iy2) : -1000, x2 = s2 >= 0 ? (TR ? iy2 : ix2) : 0;
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
  // per-seed activity: bit c * WD + w of act1 (act2) marks the cells where
  // region 1 (2) may grow at the next level
  uint32_t act1 = 0, act2 = 0;
#pragma unroll
  for (int c = 0; c < RB; c++)
#pragma unroll
    for (int w = 0; w < WD; w++) {
      if (__ballot(V1[c][w] != 0) != 0) act1 |= 1u << (c * WD + w);
      if (__ballot(V2[c][w] != 0) != 0) act2 |= 1u << (c * WD + w);
    }
  act1 = spread(act1);
  act2 = spread(act2);
  for (int round = 0;; round++) {
    uint32_t grew = 0;
#pragma unroll 1
    for (int j = 0; j < kHalo; j++) {
      uint32_t ch1 = 0, ch2 = 0;
#pragma unroll
      for (int c = 0; c < RB; c++) {
        // words in order, each updated as soon as it is computed: the left
        // neighbour's value before this level is kept in p1 / p2, the right
        // neighbour is not updated yet
        uint64_t p1 = 0, p2 = 0;
#pragma unroll
        for (int w = 0; w < WD; w++) {
          const uint32_t bit = 1u << (c * WD + w);
          const uint64_t o1 = V1[c][w], o2 = V2[c][w];
          if ((act1 | act2) & bit) {  // wave-uniform
            const uint64_t m = STSP_MASK(c, w) & ~(o1 | o2);
            uint64_t n1 = 0, n2 = 0;
            if (act1 & bit) {  // wave-uniform
              uint64_t h = (o1 << 1) | (o1 >> 1) | (p1 >> 63);
              if (w + 1 < WD) h |= V1[c][w + 1] << 63;
              n1 = (h | lane_up(o1, 0ull) | lane_down(o1, 0ull)) & m;
            }
            if (act2 & bit) {
              uint64_t h = (o2 << 1) | (o2 >> 1) | (p2 >> 63);
              if (w + 1 < WD) h |= V2[c][w + 1] << 63;
              n2 = (h | lane_up(o2, 0ull) | lane_down(o2, 0ull)) & m & ~n1;  // ties stay with seed 1
            }
            V1[c][w] = o1 | n1;
            V2[c][w] = o2 | n2;
            if (__ballot(n1 != 0) != 0) ch1 |= bit;
            if (__ballot(n2 != 0) != 0) ch2 |= bit;
            if (__ballot(own && (n1 | n2) != 0) != 0) grew = 1;
          }
          p1 = o1;
          p2 = o2;
        }
      }
      act1 = spread(ch1);
      act2 = spread(ch2);
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
    uint32_t ch1 = 0, ch2 = 0;
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
        if (__ballot(n1 != V1[c][w]) != 0) ch1 |= 1u << (c * WD + w);
        if (__ballot(n2 != V2[c][w]) != 0) ch2 |= 1u << (c * WD + w);
        V1[c][w] = n1;
        V2[c][w] = n2;
      }
    }
    // cells that changed at the last level or took new halo rows (a halo row
    // that changed may let either region grow into the cell's own rows)
    act1 |= spread(ch1 | ch2);
    act2 |= spread(ch1 | ch2);
  }
  // Seed 1's and seed 2's regions are adjacent exactly when the seeds lie in
  // one 4-connected piece of the label (a piece split in two parts has a
  // pair of adjacent pixels across the parts).  At the fixed point the halo
  // rows equal the neighbouring bands' rows, so own rows see every neighbour.
  {
    bool adj = false;
#pragma unroll
    for (int c = 0; c < RB; c++)
#pragma unroll
      for (int w = 0; w < WD; w++) {
        uint64_t h = (V1[c][w] << 1) | (V1[c][w] >> 1);
        if (w > 0) h |= V1[c][w - 1] >> 63;
        if (w + 1 < WD) h |= V1[c][w + 1] << 63;
        h |= lane_up(V1[c][w], 0ull) | lane_down(V1[c][w], 0ull);
        adj = adj || (own && (h & V2[c][w]) != 0);
      }
    if (tid == 0) E.same = 0;
    __syncthreads();
    if (__ballot(adj) != 0 && lane == 0) E.same = 1;
    __syncthreads();
  }
  const bool same = __builtin_amdgcn_readfirstlane(E.same) != 0;
  if (!TR && esc_sides) {
    // does seed 1's component (region 1, with region 2 when the seeds share
    // a piece) touch an open window edge?  Loads outside the window stay
    // inside the label's box (esc_sides marks only such edges)
    bool esc = false;
#pragma unroll
    for (int c = 0; c < RB; c++) {
      const int y = kOwn * (NW * c + wv) - kHalo + lane;
      const bool mine = own && y < bh;
      if ((esc_sides & kEscLeft) && mine && (((same ? V1[c][0] | V2[c][0] : V1[c][0])) & 1ull))
        esc = esc || seg[(by0 + y) * W + bx0 - 1] == k;
      if ((esc_sides & kEscRight) && mine) {
        const int wl = (bw - 1) >> 6;
        uint64_t word = same ? V1[c][0] | V2[c][0] : V1[c][0];
#pragma unroll
        for (int w = 1; w < WD; w++) word = w == wl ? (same ? V1[c][w] | V2[c][w] : V1[c][w]) : word;
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
          const uint64_t inside = lane_get(same ? V1[c][w] | V2[c][w] : V1[c][w], ln);
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
      // the moving part: region 2 when the seeds share a piece, else all of
      // seed 1's piece (seed 2 never reaches it)
      const uint64_t K = same ? V2[c][w] : V1[c][w];
      if (__ballot(own && K != 0) == 0) continue;  // wave-uniform: nothing moves here
      if (TR) {
        // this lane's own column, rows 64 w .. 64 w + 63: the lanes of one
        // store write neighbouring columns of one row
        const int x = row0 + lane;
        const bool mine = own && x < bw;
        for (int r = 0; r < 64; r++) {
          const int y = 64 * w + r;
          if (y >= bh) break;  // wave-uniform
          if (mine && ((K >> r) & 1ull)) halves[(by0 + y) * W + bx0 + x] = k2;
        }
        continue;
      }
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

// One workgroup of NW waves per label of work list LIST (the big-box list:
// NW = 8; a box that fits the bands only transposed is searched so, whole;
// else a window around seed 1 when the box exceeds the bands, then the
// whole-box fallbacks).
template <int NW, int LIST, int MAXR, int MAXC, class Edge>
__device__ __forceinline__ void bfs_block_items(const Batch& b, int cur, int direction, Edge& edge,
                                                uint64_t* sM, int& item) {
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
    if (!(bh <= win_h && bw <= win_w) && bw <= win_h && bh <= win_w) {
      // too wide for the bands, but it fits them transposed: searched whole
      // (box columns as band rows), no window and no fallback
      static_assert(NW * kOwn >= MAXR, "one band per wave covers the window's rows");
#define STSP_BANDS_T(WD)                                                                     \
  (void)bfs_bands<NW, 1, WD, true, Edge, true>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, \
                                               edge, 0, sM)
      if (bh <= 256) STSP_BANDS_T(4);
      else STSP_BANDS_T(8);
#undef STSP_BANDS_T
      if (tid == 0) atomicAdd(F.search_count + kSearchBand, 1ull);
      continue;
    }
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
#define STSP_BANDS(WD)                                                                          \
  escaped = bfs_bands<NW, 1, WD, true>(seg, W, k, wx0, wy0, ww, wh, s1, t2, k2, halves, edge, sides, \
                                       sM)
    if (ww <= 256) STSP_BANDS(4);
    else STSP_BANDS(8);
#undef STSP_BANDS
    if (tid == 0) {
      atomicAdd(F.search_count + (escaped ? kSearchWhole : sides ? kSearchWindow : kSearchBand), 1ull);
      // the whole-box search of an escaped window runs in kb_bfs_fallback
      if (escaped) {
        const int pos = atomicAdd(b.work_n + kWorkEsc, 1);
        if (pos < b.work_cap) b.work[3 * b.work_cap + pos] = wv;
      }
    }
  }
}

// The whole-box searches of the labels whose window search escaped (the
// escape list, work[3 work_cap..)): with LDS bitsets up to kBlkWords words,
// LDS bitsets with the words owned by threads up to kBigWords, else by
// per-pixel distances in global memory.  NW waves per workgroup.
template <int NW>
__device__ __forceinline__ void bfs_fallback_items(const Batch& b, int cur, int direction,
                                                   uint64_t* sM, uint64_t* sV1, uint64_t* sV2,
                                                   uint64_t* sNK, int* changed, int (*fb)[4],
                                                   int& item) {
  constexpr int NT = 64 * NW;
  const int tid = threadIdx.x;
  const int n_items = __builtin_amdgcn_readfirstlane(min(b.work_n[kWorkEsc], b.work_cap));
  while (true) {
    __syncthreads();
    if (tid == 0) item = atomicAdd(b.work_n + kWorkEsc + 1, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);
    if (it >= n_items) break;
    const int wv = __builtin_amdgcn_readfirstlane(b.work[3 * b.work_cap + it]);
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
    const int nwords = bh * ((bw + 63) >> 6);
    if (nwords <= kBlkWords)
      bfs_lds<NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, sM, sV1, sV2, sNK, changed);
    else if (nwords <= kBigWords)  // sM.. are one pool of 2 kBigWords here
      bfs_lds_owned<NT, kBigWords / NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, sM,
                                        sM + kBigWords, changed);
    else
      bfs_global<NT>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, halves, F.d1, F.d2, fb);
  }
}

// Boxes of the block list (up to 448 x 384 pixels, or 384 x 448, within the
// frame's search window): one workgroup of 8 waves per label, one band of 56
// rows per wave, the mask in registers.  The box is searched in the
// orientation that needs fewer 64-bit words per band row (transposed on a
// tie: its loads and stores need no cross-lane exchange); it fits the bands
// whole, so no window or fallback search is needed here.
typedef BandEdge<8, 6, 8> MidEdge;
__global__ void __launch_bounds__(512) kb_bfs_mid(Batch b, int cur, int direction) {
  __shared__ MidEdge edge;
  __shared__ int item;
  const int tid = threadIdx.x;
  // larger boxes from the list's start, then the others from its end (both
  // parts together never exceed the list: one entry per label)
  const int n_heavy = __builtin_amdgcn_readfirstlane(min(b.work_n[2], b.work_cap));
  const int n_items = __builtin_amdgcn_readfirstlane(
      min(n_heavy + min(b.work_n[kWorkMidLight], b.work_cap), b.work_cap));
  while (true) {
    __syncthreads();
    if (tid == 0) item = atomicAdd(b.work_n + 3, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item);
    if (it >= n_items) break;
    const int slot = it < n_heavy ? b.work_cap + it : 2 * b.work_cap - 1 - (it - n_heavy);
    const int wv = __builtin_amdgcn_readfirstlane(b.work[slot]);
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
    if (s1 < 0) continue;  // no pixel has d1 > d2: the label keeps all its pixels
    const bool fit_n = bh <= kMidRows && bw <= kMidCols;
    const bool fit_t = bw <= kMidRows && bh <= kMidCols;
    const int wd_n = (bw + 63) >> 6, wd_t = (bh + 63) >> 6;
    const bool tr = !fit_n || (fit_t && wd_t <= wd_n);
    const int wd = tr ? wd_t : wd_n;
#define STSP_MID(WD, T)                                                                          \
  (void)bfs_bands<8, 1, WD, false, MidEdge, T>(seg, W, k, bx0, by0, bw, bh, s1, s2, k2, F.halves, \
                                                edge, 0, nullptr)
    if (tr) {
      switch (wd) {
        case 1: STSP_MID(1, true); break;
        case 2: STSP_MID(2, true); break;
        case 3: STSP_MID(3, true); break;
        case 4: STSP_MID(4, true); break;
        case 5: STSP_MID(5, true); break;
        default: STSP_MID(6, true); break;
      }
    } else {
      switch (wd) {
        case 1: STSP_MID(1, false); break;
        case 2: STSP_MID(2, false); break;
        case 3: STSP_MID(3, false); break;
        case 4: STSP_MID(4, false); break;
        case 5: STSP_MID(5, false); break;
        default: STSP_MID(6, false); break;
      }
    }
#undef STSP_MID
    if (tid == 0) atomicAdd(F.search_count + kSearchBand, 1ull);
  }
}

// The whole-box searches' LDS: the bitsets of up to kBigWords words (two
// pools of kBigWords, or four of kBlkWords).
constexpr int kBigPool = 2 * kBigWords;
static_assert(4 * kBlkWords <= kBigPool, "four bitsets of kBlkWords fit the pool");
__global__ void __launch_bounds__(1024) kb_bfs_big(Batch b, int cur, int direction) {
  __shared__ uint64_t pool[16 * 8 * 64];  // the band mask
  __shared__ BigEdge edge;
  __shared__ int item;
  bfs_block_items<16, 2, kBigRows, kBigCols>(b, cur, direction, edge, pool, item);
}

__global__ void __launch_bounds__(512) kb_bfs_fallback(Batch b, int cur, int direction) {
  __shared__ uint64_t pool[kBigPool];
  __shared__ int item, changed[2], fb[3][4];
  bfs_fallback_items<8>(b, cur, direction, pool, pool + kBlkWords, pool + 2 * kBlkWords,
                        pool + 3 * kBlkWords, changed, fb, item);
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
  // the part that stays: the label's statistics (current before the split:
  // the sweeps keep the sums, the pass that found the boxes gathered the
  // invalid counts) less the part that moves (kb_stats_accum over the
  // halves map; zero where nothing moves)
  if (k > ms) return;  // no pixels; sst[k] of such k is a moving part, read below
  sst.count[k] = st.count[k] - sst.count[s];
  sst.ninvalid[k] = st.ninvalid[k] - sst.ninvalid[s];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sst.S[3 * k + c] = st.S[3 * k + c] - sst.S[3 * s + c];
    sst.Q[3 * k + c] = st.Q[3 * k + c] - sst.Q[3 * s + c];
  }
#pragma unroll
  for (int c = 0; c < 5; c++) sst.X[5 * k + c] = st.X[5 * k + c] - sst.X[5 * s + c];
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
  const int ms = *F.max_spix;
  if (ms > F.nbuf / 2) return;
  int* seg = F.seg[cur];
  int k[kPixPerThread];
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    const int t = t0 + j * kBlock;
    k[j] = t < F.npix ? seg[t] : -1;
  }
#pragma unroll
  for (int j = 0; j < kPixPerThread; j++) {
    if (k[j] < 0 || k[j] >= F.nbuf || !F.sm.accept[k[j]]) continue;
    const int t = t0 + j * kBlock;
    const int k2 = k[j] + (ms + 1);
    const int h = F.halves[t];
    const bool take = F.prop_mode ? (F.sm.select[k[j]] ? (h == k2) : (h == k[j])) : (h == k2);
    if (take) {
      const int sl = F.sm.pairs[2 * k[j]];
      seg[t] = sl;
      F.P.valid[sl] = 1;
    }
  }
}

// Label statistics after a split without a recount: the pass over the
// halves map and kb_split_eval left both parts of every label (k and
// k + max_spix + 1) as exact integer sums, and the labels that did not split
// kept their pixels, so an accepted label takes the sums of the part that
// stays and its new label those of the part that moves (the assignment of
// kb_split_apply).  Must run before kb_add_max changes max_spix.  (The
// invalid counts are not kept for live labels.)  The parameters of the two
// labels follow.
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
  // only the two labels of an accepted split changed statistics or priors
  // (kb_split_assign, kb_split_apply): their parameters are updated here
  param_update_one(F, k, 0);
  param_update_one(F, s, 0);
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
// Expects the halves map to equal the labels and the label statistics to
// hold the invalid counts: the statistics pass with bounding boxes that
// precedes every split writes them.
void b_split(const Batch& b, int cur, int direction, hipStream_t s, const SideStream& side) {
  b_sm_begin(b, s);
  kb_split_list<<<b.lab_blocks, kBlock, 0, s>>>(b);
  if (side.st) {
    (void)hipEventRecord(side.fork, s);
    (void)hipStreamWaitEvent(side.st, side.fork, 0);
    (void)hipStreamWaitEvent(side.st2, side.fork, 0);
  }
  kb_bfs_big<<<64, 1024, 0, side.st ? side.st2 : s>>>(b, cur, direction);
  kb_bfs_fallback<<<16, 512, 0, side.st ? side.st2 : s>>>(b, cur, direction);
  kb_bfs_mid<<<256, 512, 0, side.st ? side.st : s>>>(b, cur, direction);
  kb_bfs_wave<<<2048, 256, 0, s>>>(b, cur, direction);
  if (side.st) {
    (void)hipEventRecord(side.join, side.st);
    (void)hipEventRecord(side.join2, side.st2);
    (void)hipStreamWaitEvent(s, side.join, 0);
    (void)hipStreamWaitEvent(s, side.join2, 0);
  }
  b_moved_stats(b, cur, s);
  kb_split_eval<<<b.lab_blocks, kBlock, 0, s>>>(b);
  b_scan(b, 0, s);
  kb_split_assign<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_split_apply<<<b.pix_blocks, kBlock, 0, s>>>(b, cur);
  kb_split_stats<<<b.lab_blocks, kBlock, 0, s>>>(b);
  kb_add_max<<<1, 64, 0, s>>>(b);
}

}  // namespace stsp
