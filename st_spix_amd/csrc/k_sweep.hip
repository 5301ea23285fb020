// Fused boundary sweeps of the batched refinement loop (update_seg.cu:29-311,
// calc_joint 194-259): S checkerboard sweeps per launch over LDS tiles, with
// the statistics moves of each tile aggregated per label.
#include "kb_dev.h"

namespace stsp {

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
  // products of the exactly converted offsets: each is the rounding of the
  // same exact integer product the integer form converts
  const float f0 = (float)d0, f1 = (float)d1;
  res = res - (f0 * f0) * r.isig[0] - (f1 * f1) * r.isig[2] - ((f0 + f0) * f1) * r.isig[1];
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
  __shared__ __align__(16) int lab[LH * LW];
  // the phase lists, and after the last phase the table of statistics moves
  __shared__ __align__(16) unsigned char pool[POOLB];
  unsigned short (*list)[LCAP] = reinterpret_cast<unsigned short (*)[LCAP]>(pool);
  __shared__ int cnt[4];
  __shared__ uint32_t lutb[8];
  __shared__ uint32_t touched[(TW * TH + 31) / 32];  // tile cells whose label moved
  // Workgroup i runs on XCD i % 8; each XCD takes a contiguous run of the
  // tiles instead (raster order inside a frame), so neighbouring tiles share
  // its L2 for their overlapping label rows, image lines and records.
  const int nbk = gridDim.x, xcd = blockIdx.x & 7, q8 = nbk >> 3, r8 = nbk & 7;
  const int bid = xcd * q8 + min(xcd, r8) + (blockIdx.x >> 3);
  const int f = frame_of(b.off.tile[SHAPE], b.nf, bid);
  const FrameDesc& F = b.fr[f];
  const int H = F.H, W = F.W;
  const float* __restrict__ img = F.img;
  const int* __restrict__ seg_in = F.seg[cur];
  int* __restrict__ seg_out = F.seg[cur ^ 1];
  const SweepRec* __restrict__ rec = F.rec;
  const float sigma2_app = F.sigma2_app, potts = F.potts;
  const int tiles_x = (W + TW - 1) / TW;
  const int tile = bid - b.off.tile[SHAPE][f];
  const int tid = threadIdx.x;
  if (tid < 8) lutb[tid] = lut.bits[tid];
  for (int i = tid; i < (TW * TH + 31) / 32; i += NT) touched[i] = 0u;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  const int x0 = tx * TW - R, y0 = ty * TH - R;
  if constexpr ((NT % LW) * 8 <= NT) {
    // all loads of the tile are issued before the first LDS store: each
    // reads a clamped in-frame address, cells outside the frame get -1.
    // A thread keeps one column and steps over rows.
    constexpr int RP = NT / LW, NL = (LH + RP - 1) / RP;
    const int c = tid % LW, r0 = tid / LW;
    const int gx = x0 + c;
    const int cx = min(max(gx, 0), W - 1);
    const bool colin = (unsigned)gx < (unsigned)W;
    int v[NL];
    if (r0 < RP) {
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int gy = y0 + r0 + k * RP;
        const int cy = min(max(gy, 0), H - 1);
        if (r0 + k * RP < LH) v[k] = seg_in[cy * W + cx];
      }
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int ly = r0 + k * RP, gy = y0 + ly;
        if (ly < LH) lab[ly * LW + c] = (colin && (unsigned)gy < (unsigned)H) ? v[k] : -1;
      }
    }
  } else {
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
      static_assert(LW % 4 == 0 && R % 4 == 0, "chunks of 4 cells are 16-byte aligned");
      // chunks of 4 cells (x = 4 cx .. 4 cx + 3 of a row) covering the region
      // rows by0 .. by0 + bh - 1 and columns R - rb .. R + TW + rb - 1
      const int by0 = R - rb, bh = TH + 2 * rb;
      const int cx0 = (R - rb) >> 2, ncx = ((R + TW + rb + 3) >> 2) - cx0;
      // A thread keeps one chunk column and steps over rows, so everything
      // that depends on x alone is computed once per sweep: for cell j of the
      // chunk, whether it lies in the frame, on the frame's edge, and how far
      // outside the tile's columns (negative inside).  A cell of phase p is
      // in the phase's region when its distance outside the tile (the larger
      // of the column and row distances) is at most rb - p.
      const int rpf = NT / ncx;
      const int cc = tid % ncx, r0 = tid / ncx;
      const int x = 4 * (cx0 + cc);
      bool colin[4], coledge[4];
      int dxo[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int lx = x + j, gx = x0 + lx;
        colin[j] = (unsigned)gx < (unsigned)W;
        coledge[j] = gx < 1 || gx >= W - 1;
        dxo[j] = max(R - lx, lx - (R + TW - 1));
      }
      if (r0 < rpf) {
        for (int ly = by0 + r0; ly < by0 + bh; ly += rpf) {
          const int li = ly * LW + x;
          const int4 c4 = *reinterpret_cast<const int4*>(lab + li);
          const int4 u4 = *reinterpret_cast<const int4*>(lab + li - LW);
          const int4 d4 = *reinterpret_cast<const int4*>(lab + li + LW);
          const int lft = lab[li - 1], rgt = lab[li + 4];
          const int C[4] = {c4.x, c4.y, c4.z, c4.w};
          const int U[4] = {u4.x, u4.y, u4.z, u4.w};
          const int D[4] = {d4.x, d4.y, d4.z, d4.w};
          const int L[4] = {lft, c4.x, c4.y, c4.z};
          const int Rt[4] = {c4.y, c4.z, c4.w, rgt};
          const int gy = y0 + ly;
          const bool rowin = (unsigned)gy < (unsigned)H;
          const bool rowedge = gy < 1 || gy >= H - 1;
          const int par = gy & 1;  // x0 + x is even: phase p = 2 (j & 1) + par
          const int dyo = max(R - ly, ly - (R + TH - 1));
          const int th[2] = {rb - par, rb - 2 - par};
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const bool diff = (C[j] != U[j]) | (C[j] != D[j]) | (C[j] != L[j]) | (C[j] != Rt[j]);
            const bool inreg = max(dxo[j], dyo) <= th[j & 1];
            if (rowin && colin[j] && (rowedge || coledge[j] || diff) && inreg) {
              const int p = ((j & 1) << 1) | par;
              const int pos = atomicAdd(&cnt[p], 1);
              list[p][pos] = (unsigned short)(li + j);
            }
          }
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

}  // namespace stsp
