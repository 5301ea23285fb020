// Frame descriptors for batched (multi-frame) launches.
//
// The refinement loop of BASS / BIST (parameter update, split/merge, relabel,
// boundary sweeps) runs as launches that cover every frame of a batch: a
// device table holds one FrameDesc per frame, and each launch maps its block
// index to a frame through prefix offsets (blocks never straddle frames, so
// wave-level run aggregation stays within one frame).  A single frame is a
// batch of one.  Per-label work of uneven size (the split BFS, the relabel
// search) is pulled by persistent workgroups from compact work lists.
#pragma once

#include "kernels.h"

namespace stsp {

struct FrameDesc {
  const float* img;
  int* seg[2];            // ping-pong label maps; `cur` selects the live one
  int* halves;
  uint8_t* border;
  const int* shifted;     // propagated frames: shifted labels (split ninvalid)
  int* d1;
  int* d2;
  int H, W, npix;
  Params P;
  LabelStats st, sst;
  SmState sm;
  SweepRec* rec;
  int nbuf, nsm, cap;
  int* max_spix;
  int* total;
  int* counter;           // per-frame list counters (dead list length)
  // split proposal: boxes up to wave_box x wave_box are searched by one wave;
  // larger boxes by a workgroup in a window of at most win_h x win_w pixels
  int wave_box, win_h, win_w;
  int shape_marg;  // add the shape term to the split ratio
  unsigned long long* search_count;  // cumulative searches per path (kSearchPaths)
  // relabel
  const float* prev_prior_mu_app;
  int nspix_prev;
  int* alive;
  int* target;
  int* want_new;
  int* rank;
  int* src_of;
  int* dead;              // dead previous ids
  // settings
  float sigma2_app, potts, split_alpha, merge_alpha, thresh_relabel, thresh_new;
  int sp_size, prop_mode;
  NigConsts nc;
};

// Split-proposal search paths counted per frame: one wave; the workgroup
// bands over the whole box; the bands in a window that held seed 1's
// component; the whole box after the component left the window.
enum { kSearchWave = 0, kSearchBand, kSearchWindow, kSearchWhole, kSearchPaths };
constexpr int kWaveBoxMax = 128;
// per-pixel launches of the refinement loop: 4 pixels per thread (the block's
// frame lookup is paid once per 1024 pixels), pixel j * kBlock + thread of
// the block's 1024
constexpr int kPixPerThread = 4;
constexpr int kPixBlock = kBlock * kPixPerThread;
// label slots of a frame: labels fill 24-bit fields (work list entries
// frame << 24 | label, the sweep's packed LDS words)
constexpr int kLabelLimit = 1 << 24;
constexpr int kSplitWinRows = 896, kSplitWinCols = 512;  // the band search's register rows
constexpr int kMidRows = 448, kMidCols = 384;  // boxes of the block list (either orientation)

// Launch shapes of the fused sweep: tile width, height, sweeps per launch,
// threads.  A remainder of fewer sweeps than a shape fuses runs shape 0.
struct SweepShape {
  int tw, th, s, nt;
};
constexpr int kSweepShapes = 6;
constexpr SweepShape kSweepShape[kSweepShapes] = {
    {32, 16, 1, 64}, {32, 32, 2, 256}, {64, 32, 4, 256},
    {64, 64, 2, 256}, {64, 32, 2, 256}, {128, 64, 2, 512}};

// prefix offsets of one batch (device), each nf+1 long
struct BatchOff {
  int* pix;      // kPixBlock-pixel blocks
  int* lab;      // 256-label blocks over cap
  int* sm;       // 256-slot blocks over 2*cap+2
  int* tile[kSweepShapes];  // sweep tiles for each launch shape
  int* st;       // statistics tiles (kStatTile x kStatTile pixels)
};

constexpr int kOffArrays = 4 + kSweepShapes;
// [2i] length and [2i + 1] cursor of list i (i < 3); [kWorkMidLight] the
// length of the block list's part filled from its end; [kWorkEsc] and
// [kWorkEsc + 1] the length and cursor of list 3, the labels whose window
// search escaped
constexpr int kWorkLists = 4, kWorkMidLight = 6, kWorkEsc = 7, kWorkCounters = 9;
constexpr int kMidHeavy = 200 * 200;  // block-list boxes searched first (pixels of box area)
constexpr int kStatTile = 32;

struct Batch {
  FrameDesc* fr;  // device table
  BatchOff off;
  int nf;
  int pix_blocks, lab_blocks, sm_blocks, tile_blocks[kSweepShapes], st_blocks;
  int* work;      // kWorkLists work lists (frame << 24 | label) of work_cap entries each
  int* work_n;    // [2i] length, [2i + 1] cursor of list i
  int work_cap;
};

// host-side block counts for one frame
inline int tiles_for_shape(int shape, int H, int W) {
  const SweepShape& q = kSweepShape[shape];
  return ((W + q.tw - 1) / q.tw) * ((H + q.th - 1) / q.th);
}

inline int stat_tiles(int H, int W) {
  return ((W + kStatTile - 1) / kStatTile) * ((H + kStatTile - 1) / kStatTile);
}

// refinement-loop launches over a batch (k_batch.hip)
void b_update_params(const Batch& b, int cur, bool bbox, hipStream_t s);
void b_prepare_split(const Batch& b, int cur, hipStream_t s);
void b_param_only(const Batch& b, hipStream_t s);
void b_param_summary(const Batch& b, int cur, hipStream_t s);
void b_store_sample(const Batch& b, hipStream_t s);
int b_sweeps(const Batch& b, int cur, int niters_seg, int shape, const SimpleLut& lut,
             hipStream_t s);
// A second stream with fork/join events: launches issued on it overlap the
// main stream's between the fork and the join (also inside a graph capture).
struct SideStream {
  hipStream_t st, st2;
  hipEvent_t fork, join, join2;
};
void b_split(const Batch& b, int cur, int direction, hipStream_t s, const SideStream& side);
void b_merge(const Batch& b, int cur, int direction, hipStream_t s);
void b_relabel(const Batch& b, int cur, hipStream_t s);
void b_count_living(const Batch& b, int slot, hipStream_t s);
// pieces of the split step that live beside the statistics kernels
void b_sm_begin(const Batch& b, hipStream_t s);               // clear split/merge state, work lists
void b_moved_stats(const Batch& b, int cur, hipStream_t s);   // statistics of the moved split parts
void b_scan(const Batch& b, int which, hipStream_t s);        // per-frame exclusive scan

}  // namespace stsp
