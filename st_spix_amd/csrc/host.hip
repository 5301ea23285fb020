// Host orchestration of BASS/BIST frames and the C ABI (stspix.h).
//
// A frame runs in three parts on the caller's stream:
//   prologue   per frame: hexagonal init (BASS) or flow shift, hole filling
//              and connectivity (propagated frames), parameter set-up;
//   refinement one batched launch sequence for all frames of a batch
//              (k_batch.hip): parameter updates, split/merge, relabel and the
//              fused boundary sweeps, optionally replayed from a hipGraph;
//   epilogue   per frame: compaction and parameter export.
// Host waits happen only where later launch sizes need a device value: the
// label count after the first compaction, the hole-filling loop and the
// final superpixel count.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "../../include/stspix.h"
#include "batch.h"

using namespace stsp;

#define HIP_TRY(x)                             \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) return STSP_ERR_HIP; \
  } while (0)

namespace {

// simple-point table from its closed form: the neighbours equal to the
// candidate label (bit i in NW,N,NE,W,E,SW,S,SE order) must form exactly one
// 8-cyclic run, include a 4-neighbour, have every corner beside an edge
// member, and not be the 7-neighbour pattern with all four edges set.
SimpleLut make_lut() {
  SimpleLut L;
  memset(&L, 0, sizeof(L));
  const int ring[8] = {0, 1, 2, 4, 7, 6, 5, 3};  // NW,N,NE,E,SE,S,SW,W
  for (int num = 0; num < 256; num++) {
    int r[8], tot = 0;
    for (int i = 0; i < 8; i++) {
      r[i] = (num >> ring[i]) & 1;
      tot += r[i];
    }
    int runs = 0;
    for (int i = 0; i < 8; i++) runs += r[i] && !r[(i + 7) % 8];
    bool c1 = runs == 1;
    bool c2 = r[1] || r[3] || r[5] || r[7];
    bool c3 = true;
    for (int i = 0; i < 8; i += 2) c3 = c3 && (!r[i] || r[(i + 7) % 8] || r[(i + 1) % 8]);
    bool c4 = !(tot == 7 && r[1] && r[3] && r[5] && r[7]);
    if (c1 && c2 && c3 && c4) L.bits[num >> 5] |= 1u << (num & 31);
  }
  return L;
}

NigConsts make_nig(const stsp_config& c) {
  NigConsts n;
  float a_0 = 1e4f;
  float i_std = sqrtf(c.sigma2_app) * 2.0f;
  float b_0 = i_std * a_0;
  n.a0 = (double)a_0;
  n.b0 = (double)b_0;
  n.logb0 = stsp_det_log(n.b0);
  n.logpi = stsp_det_log(3.14159265358979323846);
  n.log2 = stsp_det_log(2.0);
  n.log_alpha = stsp_det_log((double)expf(c.alpha_hastings));
  return n;
}

struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  template <typename T>
  T* take(size_t n) {
    size_t bytes = ((sizeof(T) * (n > 0 ? n : 1)) + 255) & ~(size_t)255;
    T* p = (T*)(base + used);
    used += bytes;
    return p;
  }
};

// Event-pair timer for one kernel class (device time of the launches it
// brackets on the launch stream).
struct KTimer {
  static const int kMax = 1024;
  hipEvent_t ev[2 * kMax];
  int created = 0, n = 0;
  double ms = 0.0;
  long long count = 0;
};

}  // namespace

// Per-frame work of a batch (prologues, compaction, export) is independent
// across frames: it is spread over up to kLanes extra streams forked from the
// caller's stream with an event and joined back with events, so the many
// small launches of different frames overlap.  All streams have the default
// priority.  With lanes == 0 everything runs on the caller's stream.
constexpr int kLanes = 4;
struct Lanes {
  hipStream_t st[kLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kLanes] = {};
  int created = 0;
  int n = 0;
};

// A refinement graph captured for one batch composition.
struct CachedGraph {
  std::vector<long long> key;
  hipGraphExec_t exec = nullptr;
  int cur = 0;        // live label buffer after the graph's launches
  long long used = 0; // last use (eviction order)
};
constexpr int kGraphCache = 8;

// Batch tables on the device plus the cached refinement graphs.  Frames of
// one call that do not share a schedule are refined in groups; group g uses
// its own slice of the tables (frames [base, base + n) of h_fr/d_fr and
// offset slice g of h_off/d_off), so a group's table copy, which runs in
// stream order after the previous group's work, never reads host memory that
// a later group has rewritten.
struct stsp_batch {
  SideStream side = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int maxf = 0;
  FrameDesc* d_fr = nullptr;
  int* d_off = nullptr;  // maxf slices of kOffArrays arrays of maxf+1
  FrameDesc* h_fr = nullptr;
  int* h_off = nullptr;
  int* d_work = nullptr;
  int* d_work_n = nullptr;
  int work_cap = 0;
  Batch b;
  bool use_graph = false;
  bool noalloc = false;  // tables carved from the static scratch buffer, no side streams
  std::vector<CachedGraph> graphs;
  long long graph_clock = 0;
  Lanes lanes;
  // the last refinement's launch settings (for stsp_batch_time_sweeps)
  SimpleLut last_lut = {};
  int last_shape = -1;
  int last_nseg = 0;  // sweeps per iteration of the last refinement
  // propagation prologue: per-frame table, block offsets, per-round hole
  // counts ([kFillMax][nf] on the device, one chunk of rounds on the host)
  ProFrame* d_pro = nullptr;
  int* d_pro_off = nullptr;
  int* d_fillres = nullptr;  // per frame: filling rounds run, stall flag
  int* d_negs = nullptr;
  int* d_hlen = nullptr;  // hole-list lengths per frame
  ProFrame* h_pro = nullptr;
  int* h_pro_off = nullptr;
  int* h_fillres = nullptr;
  int* h_negs = nullptr;
  // batched epilogue: per-frame table, label/pixel block offsets (maxf + 1
  // each) and the new-label counts
  EpiFrame* d_epi = nullptr;
  int* d_epi_off = nullptr;
  int* d_epi_counts = nullptr;
  EpiFrame* h_epi = nullptr;
  int* h_epi_off = nullptr;
  int* h_epi_counts = nullptr;
};

struct stsp_ctx {
  stsp_config cfg;
  int H, W, npix;
  int cap = 0, nprev = 0;
  int sweep_shape = 4;  // index into kSweepShape (64x32 tiles, 2 sweeps per launch)
  int timing = 0;
  KTimer timers[STSP_KCLASS_COUNT];
  SimpleLut lut;
  // pixel buffers
  int *seg, *seg2, *halves, *d1, *d2, *shifted, *parent, *rsize;
  uint8_t* border;
  unsigned long long* keys;
  // label buffers
  Params P;
  LabelStats st, sst;
  SmState sm;
  SweepRec* rec;
  int *alive, *flags, *rank, *target, *want_new, *src_of, *dead;
  void* ctmp;
  // previous-frame buffers
  int64_t* fsum;
  int *counts_prev, *counts_s, *counts_f;
  float* flow_ds;
  double* centers;
  unsigned long long* best;
  // scalars
  int *max_spix, *total, *num_new, *counter;
  unsigned long long* search_count;  // kSearchPaths cumulative counters
  int wave_box = kWaveBoxMax, win_h = 0, win_w = 0;  // split search limits (0: widest)
  int shape_marginal = 0;  // the shape term in the split ratio (an extension)
  // workspace carved from the static scratch buffer: nothing is allocated,
  // the side streams are not used and the capacity is fixed
  bool noalloc = false;
  // recorded on the caller's stream at the end of every call on this
  // context: replacing a buffer waits for it rather than for the device
  hipEvent_t done_ev = nullptr;
  bool done_rec = false;
  size_t pix_bytes = 0, lab_bytes = 0, prev_bytes = 0;
  int* host;  // pinned
  void* pix_mem = nullptr;
  void* lab_mem = nullptr;
  void* prev_mem = nullptr;
  // per-frame state between prologue and epilogue
  int nbuf = 0, nspix_prev = 0, prop_mode = 0, niters = 0;
  // split/merge alphas and the new-label threshold of the current frame (the
  // label-count control changes them between iterations)
  float split_alpha = 0.f, merge_alpha = 0.f, thresh_new = 0.f;
  const float* img = nullptr;
  const float* prev_prior_mu_app = nullptr;
  stsp_batch self;  // batch of one for single-frame calls
};

// ---------------------------------------------------------------------------
// allocation
// ---------------------------------------------------------------------------
static size_t pixel_bytes(size_t n) { return 16 * 256 + n * (4 * 8 + 1 + 8) + 4096; }

// the per-pixel buffers and device scalars of a context, from `base`
static void carve_pixels(stsp_ctx* c, char* base) {
  size_t n = c->npix;
  Arena a;
  a.size = pixel_bytes(n);
  a.base = base;
  c->seg = a.take<int>(n);
  c->seg2 = a.take<int>(n);
  c->halves = a.take<int>(n);
  c->d1 = a.take<int>(n);
  c->d2 = a.take<int>(n);
  c->shifted = a.take<int>(n);
  c->parent = a.take<int>(n);
  c->rsize = a.take<int>(n);
  c->border = a.take<uint8_t>(n);
  c->keys = a.take<unsigned long long>(n);
  c->max_spix = a.take<int>(1);
  c->total = a.take<int>(1);
  c->num_new = a.take<int>(1);
  c->counter = a.take<int>(4);
  c->search_count = a.take<unsigned long long>(kSearchPaths);
}

static int alloc_pixels(stsp_ctx* c) {
  c->pix_bytes = pixel_bytes(c->npix);
  HIP_TRY(hipMalloc(&c->pix_mem, c->pix_bytes));
  carve_pixels(c, (char*)c->pix_mem);
  HIP_TRY(hipHostMalloc((void**)&c->host, 16 * sizeof(int), hipHostMallocDefault));
  return STSP_OK;
}

// Drops every cached graph.  Callers make sure none is in flight (after a
// device or stream synchronisation, or before any was launched).
static void invalidate_graph(stsp_batch* bh) {
  for (CachedGraph& g : bh->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  bh->graphs.clear();
}

static size_t label_bytes(int cap) {
  size_t n = cap, nsm = 2 * n + 2;
  return n * (3 * 4 * 2 + 8 * (2 + 3 + 2 + 3 + 3 + 3 + 1) + 4 * 6) + n * (4 + 8 * 11 + 4 + 16) +
         nsm * (4 + 8 * 11 + 4) + nsm * (4 * 4 + 8 * 3 * 2 + 8 * 2 + 4 * 8) + nsm * 8 +
         n * sizeof(SweepRec) + n * 4 * 8 + compact_tmp_bytes(cap) + 100 * 256;
}

// the per-label workspace of `cap` slots, from `base`
static int carve_labels(stsp_ctx* c, char* base, int cap) {
  size_t n = cap, nsm = 2 * n + 2;
  Arena a;
  a.base = base;
  a.size = label_bytes(cap);
  Params& P = c->P;
  P.mu_app = a.take<float>(3 * n);
  P.prior_mu_app = a.take<float>(3 * n);
  P.mu_shape = a.take<double>(2 * n);
  P.sigma_shape = a.take<double>(3 * n);
  P.prior_mu_shape = a.take<double>(2 * n);
  P.prior_sigma_shape = a.take<double>(3 * n);
  P.sample_sigma_shape = a.take<double>(3 * n);
  P.prior_icov = a.take<double>(3 * n);
  P.logdet = a.take<double>(n);
  P.count = a.take<int>(n);
  P.icount = a.take<float>(n);
  P.sm_count = a.take<int>(n);
  P.prior_count = a.take<float>(n);
  P.valid = a.take<int>(n);
  P.prop = a.take<int>(n);
  c->st.count = a.take<int>(n);
  c->st.S = a.take<int64_t>(3 * n);
  c->st.Q = a.take<int64_t>(3 * n);
  c->st.X = a.take<int64_t>(5 * n);
  c->st.ninvalid = a.take<int>(n);
  c->st.bbox = a.take<int>(4 * n);
  c->sst.count = a.take<int>(nsm);
  c->sst.S = a.take<int64_t>(3 * nsm);
  c->sst.Q = a.take<int64_t>(3 * nsm);
  c->sst.X = a.take<int64_t>(5 * nsm);
  c->sst.ninvalid = a.take<int>(nsm);
  c->sst.bbox = nullptr;
  SmState& sm = c->sm;
  sm.count = c->sst.count;
  sm.S = c->sst.S;
  sm.Q = c->sst.Q;
  sm.ninvalid = c->sst.ninvalid;
  sm.count_f = a.take<int>(nsm);
  sm.b_n = a.take<double>(3 * nsm);
  sm.b_n_f = a.take<double>(3 * nsm);
  sm.marg = a.take<double>(nsm);
  sm.hasting = a.take<double>(nsm);
  sm.merge = a.take<int>(nsm);
  sm.remove = a.take<int>(nsm);
  sm.select = a.take<int>(nsm);
  sm.has_bn = a.take<int>(nsm);
  sm.pairs = a.take<int>(2 * nsm);
  sm.winner = a.take<int>(nsm);
  sm.accept = a.take<int>(nsm);
  sm.rank = a.take<int>(nsm);
  sm.valid0 = a.take<int>(nsm);
  sm.pc0 = a.take<float>(nsm);
  c->rec = a.take<SweepRec>(n);
  c->alive = a.take<int>(n);
  c->flags = a.take<int>(n);
  c->rank = a.take<int>(n);
  c->target = a.take<int>(n);
  c->want_new = a.take<int>(n);
  c->src_of = a.take<int>(n);
  c->dead = a.take<int>(n);
  c->ctmp = (void*)a.take<char>(compact_tmp_bytes(cap));
  if (a.used > a.size) {
    fprintf(stderr, "stspix: label arena overflow (%zu > %zu)\n", a.used, a.size);
    return STSP_ERR_HIP;
  }
  c->cap = cap;
  return STSP_OK;
}

// (re)allocate the per-label workspace for `cap` slots
static int ensure_labels(stsp_ctx* c, int cap) {
  if (cap <= c->cap) return STSP_OK;
  if (c->noalloc) return STSP_ERR_TOO_LARGE;  // sized once, from the static scratch buffer
  if (cap + cap / 4 > kLabelLimit) return STSP_ERR_TOO_LARGE;
  if (c->lab_mem) {
    if (c->done_rec) HIP_TRY(hipEventSynchronize(c->done_ev));  // the last call's work
    HIP_TRY(hipFree(c->lab_mem));
    c->lab_mem = nullptr;
    c->lab_bytes = 0;
  }
  cap = cap + cap / 4;  // headroom for growth across frames
  HIP_TRY(hipMalloc(&c->lab_mem, label_bytes(cap)));
  c->lab_bytes = label_bytes(cap);
  const int rc = carve_labels(c, (char*)c->lab_mem, cap);
  invalidate_graph(&c->self);
  return rc;
}

static int ensure_prev(stsp_ctx* c, int n) {
  if (n <= c->nprev) return STSP_OK;
  if (c->prev_mem) {
    if (c->done_rec) HIP_TRY(hipEventSynchronize(c->done_ev));  // the last call's work
    HIP_TRY(hipFree(c->prev_mem));
    c->prev_mem = nullptr;
    c->prev_bytes = 0;
  }
  n = n + n / 4;
  size_t bytes = (size_t)n * (16 + 12 + 8 + 16 + 8) + 16 * 256;
  HIP_TRY(hipMalloc(&c->prev_mem, bytes));
  c->prev_bytes = bytes;
  Arena a;
  a.base = (char*)c->prev_mem;
  a.size = bytes;
  c->fsum = a.take<int64_t>(2 * (size_t)n);
  c->counts_prev = a.take<int>(n);
  c->counts_s = a.take<int>(n);
  c->counts_f = a.take<int>(n);
  c->flow_ds = a.take<float>(2 * (size_t)n);
  c->centers = a.take<double>(2 * (size_t)n);
  c->best = a.take<unsigned long long>(n);
  c->nprev = n;
  return STSP_OK;
}

static int batch_alloc(stsp_batch* bh, int maxf) {
  bh->maxf = maxf;
  HIP_TRY(hipMalloc((void**)&bh->d_fr, sizeof(FrameDesc) * maxf));
  const size_t off_bytes = sizeof(int) * kOffArrays * (maxf + 1) * maxf;
  HIP_TRY(hipMalloc((void**)&bh->d_off, off_bytes));
  HIP_TRY(hipMalloc((void**)&bh->d_work_n, sizeof(int) * kWorkCounters));
  HIP_TRY(hipHostMalloc((void**)&bh->h_fr, sizeof(FrameDesc) * maxf, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_off, off_bytes, hipHostMallocDefault));
  HIP_TRY(hipMalloc((void**)&bh->d_pro, sizeof(ProFrame) * maxf));
  HIP_TRY(hipMalloc((void**)&bh->d_pro_off, sizeof(int) * kProKinds * (kProMax + 1)));
  HIP_TRY(hipMalloc((void**)&bh->d_fillres, sizeof(int) * 2 * maxf));
  HIP_TRY(hipMalloc((void**)&bh->d_negs, sizeof(int) * maxf * kNegSub));
  HIP_TRY(hipMalloc((void**)&bh->d_hlen, sizeof(int) * maxf * kHoleSub));
  HIP_TRY(hipHostMalloc((void**)&bh->h_pro, sizeof(ProFrame) * maxf, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_pro_off, sizeof(int) * kProKinds * (kProMax + 1),
                        hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_fillres, sizeof(int) * 2 * maxf,
                        hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_negs, sizeof(int) * maxf * kNegSub,
                        hipHostMallocDefault));
  HIP_TRY(hipMalloc((void**)&bh->d_epi, sizeof(EpiFrame) * maxf));
  HIP_TRY(hipMalloc((void**)&bh->d_epi_off, sizeof(int) * 2 * (maxf + 1)));
  HIP_TRY(hipMalloc((void**)&bh->d_epi_counts, sizeof(int) * maxf));
  HIP_TRY(hipHostMalloc((void**)&bh->h_epi, sizeof(EpiFrame) * maxf, hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_epi_off, sizeof(int) * 2 * (maxf + 1),
                        hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&bh->h_epi_counts, sizeof(int) * maxf, hipHostMallocDefault));
  return STSP_OK;
}

static int lanes_fork(Lanes* L, hipStream_t s, int nframes) {
  if (!L) return STSP_OK;
  L->n = 0;
  if (nframes < 2) return STSP_OK;
  if (!L->created) {
    for (int i = 0; i < kLanes; i++) {
      HIP_TRY(hipStreamCreateWithFlags(&L->st[i], hipStreamNonBlocking));
      HIP_TRY(hipEventCreateWithFlags(&L->join[i], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&L->fork, hipEventDisableTiming));
    L->created = 1;
  }
  L->n = nframes < kLanes ? nframes : kLanes;
  HIP_TRY(hipEventRecord(L->fork, s));
  for (int i = 0; i < L->n; i++) HIP_TRY(hipStreamWaitEvent(L->st[i], L->fork, 0));
  return STSP_OK;
}

static inline hipStream_t lane_of(const Lanes* L, hipStream_t s, int frame) {
  return (L && L->n) ? L->st[frame % L->n] : s;
}

// host wait for everything issued on the lanes (or on s without lanes)
static int lanes_sync(const Lanes* L, hipStream_t s) {
  if (!L || !L->n) {
    HIP_TRY(hipStreamSynchronize(s));
    return STSP_OK;
  }
  for (int i = 0; i < L->n; i++) HIP_TRY(hipStreamSynchronize(L->st[i]));
  return STSP_OK;
}

static int lanes_join(Lanes* L, hipStream_t s) {
  if (!L || !L->n) return STSP_OK;
  for (int i = 0; i < L->n; i++) {
    HIP_TRY(hipEventRecord(L->join[i], L->st[i]));
    HIP_TRY(hipStreamWaitEvent(s, L->join[i], 0));
  }
  L->n = 0;
  return STSP_OK;
}

static void lanes_free(Lanes* L) {
  if (!L->created) return;
  for (int i = 0; i < kLanes; i++) {
    (void)hipStreamSynchronize(L->st[i]);
    (void)hipStreamDestroy(L->st[i]);
    (void)hipEventDestroy(L->join[i]);
  }
  (void)hipEventDestroy(L->fork);
  L->created = 0;
}

static void batch_free(stsp_batch* bh) {
  lanes_free(&bh->lanes);
  if (bh->side.st) {
    (void)hipStreamSynchronize(bh->side.st);
    (void)hipStreamSynchronize(bh->side.st2);
    (void)hipStreamDestroy(bh->side.st);
    (void)hipStreamDestroy(bh->side.st2);
    (void)hipEventDestroy(bh->side.fork);
    (void)hipEventDestroy(bh->side.join);
    (void)hipEventDestroy(bh->side.join2);
    bh->side = {nullptr, nullptr, nullptr, nullptr, nullptr};
  }
  invalidate_graph(bh);
  if (bh->d_fr) (void)hipFree(bh->d_fr);
  if (bh->d_off) (void)hipFree(bh->d_off);
  if (bh->d_work) (void)hipFree(bh->d_work);
  if (bh->d_work_n) (void)hipFree(bh->d_work_n);
  if (bh->h_fr) (void)hipHostFree(bh->h_fr);
  if (bh->h_off) (void)hipHostFree(bh->h_off);
  if (bh->d_pro) (void)hipFree(bh->d_pro);
  if (bh->d_pro_off) (void)hipFree(bh->d_pro_off);
  if (bh->d_fillres) (void)hipFree(bh->d_fillres);
  if (bh->d_negs) (void)hipFree(bh->d_negs);
  if (bh->d_hlen) (void)hipFree(bh->d_hlen);
  if (bh->h_pro) (void)hipHostFree(bh->h_pro);
  if (bh->h_pro_off) (void)hipHostFree(bh->h_pro_off);
  if (bh->h_fillres) (void)hipHostFree(bh->h_fillres);
  if (bh->h_negs) (void)hipHostFree(bh->h_negs);
  if (bh->d_epi) (void)hipFree(bh->d_epi);
  if (bh->d_epi_off) (void)hipFree(bh->d_epi_off);
  if (bh->d_epi_counts) (void)hipFree(bh->d_epi_counts);
  if (bh->h_epi) (void)hipHostFree(bh->h_epi);
  if (bh->h_epi_off) (void)hipHostFree(bh->h_epi_off);
  if (bh->h_epi_counts) (void)hipHostFree(bh->h_epi_counts);
  bh->d_epi = nullptr;
  bh->h_epi = nullptr;
  bh->d_epi_off = bh->d_epi_counts = bh->h_epi_off = bh->h_epi_counts = nullptr;
  bh->d_pro = nullptr;
  bh->d_pro_off = nullptr;
  bh->d_fillres = bh->d_negs = bh->d_hlen = nullptr;
  bh->h_pro = nullptr;
  bh->h_pro_off = nullptr;
  bh->h_fillres = bh->h_negs = nullptr;
  bh->d_fr = nullptr;
  bh->d_off = nullptr;
  bh->d_work = nullptr;
  bh->d_work_n = nullptr;
  bh->h_fr = nullptr;
  bh->h_off = nullptr;
  bh->work_cap = 0;
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
static void tstart(stsp_ctx* c, int cls, hipStream_t s) {
  if (!c || !c->timing) return;
  KTimer& t = c->timers[cls];
  if (t.n >= KTimer::kMax) return;
  if (t.created <= t.n) {
    // timing-only events: no system-scope fence (cache write-back and
    // invalidation) at record time, which would be counted in the interval
    (void)hipEventCreateWithFlags(&t.ev[2 * t.n], hipEventDisableSystemFence);
    (void)hipEventCreateWithFlags(&t.ev[2 * t.n + 1], hipEventDisableSystemFence);
    t.created = t.n + 1;
  }
  (void)hipEventRecord(t.ev[2 * t.n], s);
}
static void tstop(stsp_ctx* c, int cls, hipStream_t s) {
  if (!c || !c->timing) return;
  KTimer& t = c->timers[cls];
  if (t.n >= KTimer::kMax) return;
  (void)hipEventRecord(t.ev[2 * t.n + 1], s);
  t.n++;
}
static void tflush(stsp_ctx* c) {
  if (!c->timing) return;
  for (int k = 0; k < STSP_KCLASS_COUNT; k++) {
    KTimer& t = c->timers[k];
    for (int i = 0; i < t.n; i++) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.ev[2 * i], t.ev[2 * i + 1]) == hipSuccess) {
        t.ms += ms;
        t.count++;
      }
    }
    t.n = 0;
  }
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static int read_int(stsp_ctx* c, const int* dev, hipStream_t s, int* out) {
  HIP_TRY(hipMemcpyAsync(c->host, dev, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *out = c->host[0];
  return STSP_OK;
}

// end of a call on context c: everything it queued is ordered before this
static int mark_done(stsp_ctx* c, hipStream_t s) {
  if (!c->done_ev) return STSP_OK;
  HIP_TRY(hipEventRecord(c->done_ev, s));
  c->done_rec = true;
  return STSP_OK;
}

// ---------------------------------------------------------------------------
// batch tables
// ---------------------------------------------------------------------------
static FrameDesc make_desc(stsp_ctx* c) {
  FrameDesc d;
  memset(&d, 0, sizeof(d));
  d.img = c->img;
  d.seg[0] = c->seg;
  d.seg[1] = c->seg2;
  d.halves = c->halves;
  d.border = c->border;
  d.shifted = c->prop_mode ? c->shifted : nullptr;
  d.d1 = c->d1;
  d.d2 = c->d2;
  d.H = c->H;
  d.W = c->W;
  d.npix = c->npix;
  d.P = c->P;
  d.st = c->st;
  d.sst = c->sst;
  d.sm = c->sm;
  d.rec = c->rec;
  d.nbuf = c->nbuf;
  d.nsm = 2 * c->nbuf + 2;
  d.cap = c->cap;
  d.max_spix = c->max_spix;
  d.total = c->total;
  d.counter = c->counter;
  d.wave_box = c->wave_box;
  d.shape_marg = c->shape_marginal;
  d.win_h = c->win_h > 0 ? c->win_h : kSplitWinRows;
  d.win_w = c->win_w > 0 ? c->win_w : kSplitWinCols;
  d.search_count = c->search_count;
  d.prev_prior_mu_app = c->prev_prior_mu_app;
  d.nspix_prev = c->nspix_prev;
  d.alive = c->alive;
  d.target = c->target;
  d.want_new = c->want_new;
  d.rank = c->rank;
  d.src_of = c->src_of;
  d.dead = c->dead;
  d.sigma2_app = c->cfg.sigma2_app;
  d.potts = c->cfg.potts;
  d.split_alpha = c->split_alpha;
  d.merge_alpha = c->merge_alpha;
  d.thresh_relabel = c->cfg.thresh_relabel;
  d.thresh_new = c->thresh_new;
  d.sp_size = c->cfg.sp_size;
  d.prop_mode = c->prop_mode;
  d.nc = make_nig(c->cfg);
  return d;
}

// Work lists (split candidates, relabel items) for `work` labels; grows only
// with the stream idle, and drops the graphs that captured the old lists.
static int ensure_work(stsp_batch* bh, int work, hipStream_t s) {
  if (work <= bh->work_cap) return STSP_OK;
  if (bh->noalloc) return STSP_ERR_TOO_LARGE;
  if (bh->d_work) {
    HIP_TRY(hipStreamSynchronize(s));
    if (bh->side.st) HIP_TRY(hipStreamSynchronize(bh->side.st));
    if (bh->side.st2) HIP_TRY(hipStreamSynchronize(bh->side.st2));
    HIP_TRY(hipFree(bh->d_work));
    bh->d_work = nullptr;
  }
  invalidate_graph(bh);
  HIP_TRY(hipMalloc((void**)&bh->d_work, kWorkLists * sizeof(int) * (size_t)work));
  bh->work_cap = work;
  return STSP_OK;
}

// Fills the device table of group `slot` for frames ctxs[0..n), stored at
// frames [base, base + n) of the frame table; grid sizes cover each frame's
// capacity so that a captured graph stays valid across frames.
static int batch_build(stsp_batch* bh, stsp_ctx** ctxs, int n, hipStream_t s, int base = 0,
                       int slot = 0) {
  if (n < 1 || base < 0 || slot < 0 || base + n > bh->maxf || slot >= bh->maxf || n > 64)
    return STSP_ERR_ARG;
  if (!bh->side.st && !bh->noalloc) {
    HIP_TRY(hipStreamCreateWithFlags(&bh->side.st, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&bh->side.st2, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&bh->side.fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&bh->side.join, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&bh->side.join2, hipEventDisableTiming));
  }
  const int M = bh->maxf + 1;
  int* off = bh->h_off + (size_t)slot * kOffArrays * M;
  FrameDesc* hfr = bh->h_fr + base;
  int* o_pix = off;
  int* o_lab = off + M;
  int* o_sm = off + 2 * M;
  int* o_t[kSweepShapes];
  for (int sh = 0; sh < kSweepShapes; sh++) o_t[sh] = off + (3 + sh) * M;
  int* o_st = off + (3 + kSweepShapes) * M;
  int work = 0;
  o_pix[0] = o_lab[0] = o_sm[0] = o_st[0] = 0;
  for (int sh = 0; sh < kSweepShapes; sh++) o_t[sh][0] = 0;
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = ctxs[i];
    hfr[i] = make_desc(c);
    o_pix[i + 1] = o_pix[i] + div_up(c->npix, kPixBlock);
    o_lab[i + 1] = o_lab[i] + div_up(c->cap, kBlock);
    o_sm[i + 1] = o_sm[i] + div_up(2 * c->cap + 2, kBlock);
    for (int sh = 0; sh < kSweepShapes; sh++)
      o_t[sh][i + 1] = o_t[sh][i] + tiles_for_shape(sh, c->H, c->W);
    o_st[i + 1] = o_st[i] + stat_tiles(c->H, c->W);
    work += c->cap;
  }
  int rc = ensure_work(bh, work, s);
  if (rc) return rc;
  int* d = bh->d_off + (size_t)slot * kOffArrays * M;
  HIP_TRY(hipMemcpyAsync(bh->d_fr + base, hfr, sizeof(FrameDesc) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(d, off, sizeof(int) * kOffArrays * M, hipMemcpyHostToDevice, s));
  Batch& b = bh->b;
  b.fr = bh->d_fr + base;
  b.off.pix = d;
  b.off.lab = d + M;
  b.off.sm = d + 2 * M;
  for (int sh = 0; sh < kSweepShapes; sh++) b.off.tile[sh] = d + (3 + sh) * M;
  b.nf = n;
  b.pix_blocks = o_pix[n];
  b.lab_blocks = o_lab[n];
  b.sm_blocks = o_sm[n];
  for (int sh = 0; sh < kSweepShapes; sh++) b.tile_blocks[sh] = o_t[sh][n];
  b.off.st = d + (3 + kSweepShapes) * M;
  b.st_blocks = o_st[n];
  b.work = bh->d_work;
  b.work_n = bh->d_work_n;
  b.work_cap = bh->work_cap;
  return STSP_OK;
}

static inline bool is_split_iter(const stsp_config& cfg, int idx) {
  int p = cfg.sm_period > 0 ? cfg.sm_period : 4;
  return idx > cfg.sm_start && idx % p == 0;
}
static inline bool is_merge_iter(const stsp_config& cfg, int idx) {
  int p = cfg.sm_period > 0 ? cfg.sm_period : 4;
  return idx > cfg.sm_start && idx % p == p / 2;
}

// One outer iteration of bass() (bass.cu:71-168) / prop_bass_v3
// (prop.cu:364-457) for every frame of the batch; the update of the
// parameters at its start is done by the caller.  `lead` supplies the
// settings shared by the batch and the timers.
static void iteration_tail(stsp_ctx* lead, const Batch& b, const SideStream& side, int idx,
                           int prop_mode, int shape, int* cur, int* smc, hipStream_t s) {
  const stsp_config& cfg = lead->cfg;
  // split and merge leave the label statistics current (kb_split_stats,
  // kb_merge_stats); the split also updates the parameters of the labels it
  // changed, the merge is followed by a parameter update
  if (is_split_iter(cfg, idx)) {
    *smc += 1;
    b_split(b, *cur, *smc % 2 + 1, s, side);
  }
  if (is_merge_iter(cfg, idx)) {
    if (prop_mode) {
      b_relabel(b, *cur, s);
      *smc += 1;
    }
    b_merge(b, *cur, *smc % 2, s);
    b_param_only(b, s);
  }
  tstart(lead, STSP_KCLASS_SWEEP, s);
  *cur = b_sweeps(b, *cur, cfg.niters_seg, shape, lead->lut, s);
  tstop(lead, STSP_KCLASS_SWEEP, s);
}

// The parameter update at the start of an iteration.  The sweeps keep the
// label statistics current (a changed pixel moves its integer sums) and
// split and merge combine them exactly, so a full recount is needed only at
// the first iteration (the prologue leaves no statistics); before a split a
// lighter pass gathers the bounding boxes.
static void update_at_start(const Batch& b, const stsp_config& cfg, int idx, int cur,
                            hipStream_t s) {
  if (idx == 0)  // bounding boxes and the halves map too if it is a split iteration
    b_update_params(b, cur, is_split_iter(cfg, idx), s);
  else if (is_split_iter(cfg, idx))
    b_prepare_split(b, cur, s);
  else
    b_param_only(b, s);
}

// The refinement loop for every frame of the batch; returns the live
// buffer index.
static int refine_launches(stsp_ctx* lead, const Batch& b, const SideStream& side, int niters,
                           int prop_mode, int shape, hipStream_t s) {
  const stsp_config& cfg = lead->cfg;
  int cur = 0;
  int smc = 1;  // split/merge counter (bass.cu:56, prop.cu:343)
  for (int idx = 0; idx < niters; idx++) {
    tstart(lead, STSP_KCLASS_STATS, s);
    update_at_start(b, cfg, idx, cur, s);
    tstop(lead, STSP_KCLASS_STATS, s);
    iteration_tail(lead, b, side, idx, prop_mode, shape, &cur, &smc, s);
  }
  b_update_params(b, cur, false, s);
  b_store_sample(b, s);
  return cur;
}

// Split/merge alpha control toward target_nspix (bass.cu:73-98,
// prop.cu:371-400; the prop variant also doubles thresh_new every 200
// iterations past 2000).  Returns true when the count is accepted.
static bool control_nspix(int nliving, int target, int idx, int og_niters, int prop_mode,
                          float* split_alpha, float* merge_alpha, float* thresh_new) {
  if (nliving > 1.05 * target) {
    if (*split_alpha > 0) {
      *split_alpha = 0;
      *merge_alpha = 0;
    }
    *split_alpha += -1;
    *merge_alpha += 1.0;
  } else if (nliving < 0.95 * target) {
    if (*split_alpha < 0) {
      *split_alpha = 0;
      *merge_alpha = 0;
    }
    *split_alpha += 1;
    *merge_alpha += -1;
  } else {
    *split_alpha = 0.;
    *merge_alpha = 0.;
  }
  if (prop_mode && idx > 2000 && idx % 200 == 0) *thresh_new = (float)(2 * (*thresh_new + 1e-5));
  return nliving < 1.05 * target && nliving > 0.95 * target && idx >= og_niters;
}

// Refinement of one frame with the label-count control: every second
// iteration the number of distinct labels (labels with pixels after the
// parameter update) is read back and the alphas are adjusted; the loop ends
// when the count is accepted or after 5000 iterations.  The reference counts
// before the update in BASS and after it in prop; the update does not change
// the labels and repeating it is idempotent, so both orders give the same
// result.
static int refine_controlled(stsp_batch* bh, stsp_ctx* c, hipStream_t s) {
  stsp_ctx* one = c;
  int rc = batch_build(bh, &one, 1, s);
  if (rc) return rc;
  const stsp_config& cfg = c->cfg;
  const int og_niters = c->niters;
  int cur = 0, smc = 1;
  for (int idx = 0; idx < 5000; idx++) {
    update_at_start(bh->b, cfg, idx, cur, s);
    if (idx % 2 == 0) {
      k_set_int<<<1, 1, 0, s>>>(c->counter + 1, 0);
      b_count_living(bh->b, 1, s);
      int nliving = 0;
      rc = read_int(c, c->counter + 1, s, &nliving);
      if (rc) return rc;
      if (control_nspix(nliving, cfg.target_nspix, idx, og_niters, c->prop_mode, &c->split_alpha,
                        &c->merge_alpha, &c->thresh_new))
        break;
      rc = batch_build(bh, &one, 1, s);  // the stream is idle here: the table can be rewritten
      if (rc) return rc;
    }
    iteration_tail(c, bh->b, bh->side, idx, c->prop_mode, c->sweep_shape, &cur, &smc, s);
  }
  b_update_params(bh->b, cur, false, s);
  b_store_sample(bh->b, s);
  if (cur == 1) {
    int* t = c->seg;
    c->seg = c->seg2;
    c->seg2 = t;
  }
  HIP_TRY(hipGetLastError());
  return STSP_OK;
}

// Runs the refinement of frames ctxs[0..n) (prologues done) as one batch,
// with the tables of group `slot` at frame `base` (see stsp_batch).
static int refine(stsp_batch* bh, stsp_ctx** ctxs, int n, hipStream_t s, int base = 0,
                  int slot = 0) {
  stsp_ctx* lead = ctxs[0];
  if (lead->cfg.target_nspix > 0) return n == 1 ? refine_controlled(bh, lead, s) : STSP_ERR_ARG;
  const int niters = lead->niters, prop_mode = lead->prop_mode, shape = lead->sweep_shape;
  for (int i = 1; i < n; i++) {
    const stsp_config& a = ctxs[i]->cfg;
    const stsp_config& l = lead->cfg;
    if (ctxs[i]->niters != niters || ctxs[i]->prop_mode != prop_mode ||
        a.niters_seg != l.niters_seg || a.sm_period != l.sm_period || a.sm_start != l.sm_start)
      return STSP_ERR_ARG;
  }
  int rc = batch_build(bh, ctxs, n, s, base, slot);
  if (rc) return rc;
  bh->last_lut = lead->lut;
  bh->last_shape = shape;
  bh->last_nseg = lead->cfg.niters_seg;
  int cur;
  bool timing = lead->timing != 0;
  // The legacy default stream cannot be captured: there the launches run
  // eagerly (same kernels, same results).
  if (bh->use_graph && !timing && s != nullptr) {
    std::vector<long long> key = {n, niters, prop_mode, shape, lead->cfg.niters_seg,
                                  lead->cfg.sm_period, lead->cfg.sm_start, bh->work_cap,
                                  base, slot};
    for (int i = 0; i < n; i++) {
      key.push_back((long long)(size_t)ctxs[i]);  // grid sizes follow from cap, H and W
      key.push_back(ctxs[i]->cap);
      key.push_back(ctxs[i]->H);
      key.push_back(ctxs[i]->W);
    }
    CachedGraph* hit = nullptr;
    for (CachedGraph& g : bh->graphs)
      if (g.key == key) hit = &g;
    if (!hit) {
      if ((int)bh->graphs.size() >= kGraphCache) {
        // evict the least recently used graph once the stream has drained
        size_t lru = 0;
        for (size_t i = 1; i < bh->graphs.size(); i++)
          if (bh->graphs[i].used < bh->graphs[lru].used) lru = i;
        HIP_TRY(hipStreamSynchronize(s));
        (void)hipGraphExecDestroy(bh->graphs[lru].exec);
        bh->graphs.erase(bh->graphs.begin() + lru);
      }
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int c0 = refine_launches(lead, bh->b, bh->side, niters, prop_mode, shape, s);
      hipGraph_t g = nullptr;
      const hipError_t le = hipGetLastError();
      const hipError_t ee = hipStreamEndCapture(s, &g);  // always leave capture mode
      if (le != hipSuccess || ee != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return STSP_ERR_HIP;
      }
      CachedGraph cg;
      const hipError_t ie = hipGraphInstantiate(&cg.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return STSP_ERR_HIP;
      cg.key = key;
      cg.cur = c0;
      bh->graphs.push_back(cg);
      hit = &bh->graphs.back();
    }
    hit->used = ++bh->graph_clock;
    HIP_TRY(hipGraphLaunch(hit->exec, s));
    cur = hit->cur;
  } else {
    cur = refine_launches(lead, bh->b, bh->side, niters, prop_mode, shape, s);
  }
  for (int i = 0; i < n; i++) {
    if (cur == 1) {
      int* t = ctxs[i]->seg;
      ctxs[i]->seg = ctxs[i]->seg2;
      ctxs[i]->seg2 = t;
    }
  }
  HIP_TRY(hipGetLastError());
  return STSP_OK;
}

// ---------------------------------------------------------------------------
// prologues and epilogue
// ---------------------------------------------------------------------------
// First frames (BASS set-up, bass.cu:178-330) of frames cs[0..n): the
// hexagonal initialisation and compaction of every frame, one wait for all
// label counts, then the parameter set-up.  Frame i runs on lane i.
static int bass_prologue_n(stsp_ctx** cs, const float* const* imgs, int n, Lanes* L,
                           hipStream_t s) {
  int rc = lanes_fork(L, s, n);
  if (rc) return rc;
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = cs[i];
    const hipStream_t ss = lane_of(L, s, i);
    const stsp_config& cfg = c->cfg;
    HexGeom g = hex_geometry(cfg.sp_size, c->W, c->H);
    int K0 = g.nx * g.ny;
    int nbuf = 30 * K0;  // bass.cu:245
    rc = ensure_labels(c, nbuf);
    if (rc) return rc;
    c->nbuf = nbuf;
    launch_hex_init(c->seg, c->H, c->W, cfg.sp_size, ss);
    launch_compact(c->seg, c->npix, nullptr, K0, 0, nbuf, c->alive, c->flags, c->rank,
                   c->num_new, c->ctmp, ss);
    HIP_TRY(hipMemcpyAsync(c->host, c->num_new, sizeof(int), hipMemcpyDeviceToHost, ss));
  }
  rc = lanes_sync(L, s);
  if (rc) return rc;
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = cs[i];
    const hipStream_t ss = lane_of(L, s, i);
    const stsp_config& cfg = c->cfg;
    const int nspix = c->host[0];
    c->img = imgs[i];
    c->nspix_prev = 0;
    c->prop_mode = 0;
    c->niters = cfg.niters;
    c->split_alpha = cfg.split_alpha;
    c->merge_alpha = 0.f;  // local to bass() (bass.cu:55)
    c->thresh_new = cfg.thresh_new;
    c->prev_prior_mu_app = nullptr;
    // mark_active_contiguous + init_sp_params (init_sparams.cu:25-231)
    launch_bass_mark(c->P, nspix, c->nbuf, cfg.sp_size, ss);
    stsp_ctx* one = c;
    rc = batch_build(&c->self, &one, 1, ss);
    if (rc) return rc;
    b_param_summary(c->self.b, 0, ss);
    launch_bass_init(c->P, nspix, c->nbuf, cfg.sp_size, c->npix, ss);
    k_set_int<<<1, 1, 0, ss>>>(c->max_spix, nspix - 1);
  }
  HIP_TRY(hipGetLastError());
  return lanes_join(L, s);
}

// Hole filling (fill_missing.cu:35-130) of every frame of the prologue: the
// holes are listed once, one launch runs all rounds of every frame (a
// workgroup per frame, kp_fill_all), and one wait reads each frame's outcome.
// A hole count that stops shrinking is an error as in the reference
// (fill_missing.cu:97-111).
static int fill_loop_pro(stsp_batch* bh, const Pro& p, const int* total, hipStream_t s) {
  pro_hole_list(p, total, s);
  pro_fill_all(p, s);
  HIP_TRY(hipMemcpyAsync(bh->h_fillres, p.fillres, sizeof(int) * 2 * p.nall,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < p.nall; i++)
    if (bh->h_fillres[2 * i + 1]) return STSP_ERR_FILL;
  return STSP_OK;
}

// shift_and_fill (shift_and_fill.cu:108-339) + run_prop set-up
// (prop.cu:677-745) for frames cs[0..n): every stage is one launch over all
// frames (the ProFrame table), and the host waits only for the hole counts.
// shifted_out / filled_out (may be null, or hold null entries) receive
// copies of each frame's shifted map (holes -1) and of the filled map that
// starts the refinement.
static int prop_prologue_n(stsp_batch* bh, stsp_ctx** cs, int n, const float* const* imgs,
                           const float* const* flows, const int* const* spix_prev,
                           stsp_params* prevs, const int* nprev, int* niters_state,
                           hipStream_t s, int* const* shifted_out = nullptr,
                           int* const* filled_out = nullptr) {
  if (n > bh->maxf || n > kProMax) return STSP_ERR_ARG;
  int* off = bh->h_pro_off;
  for (int k = 0; k < kProKinds; k++) off[k * (kProMax + 1)] = 0;
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = cs[i];
    const int np = nprev[i];
    int rc = ensure_labels(c, 10 * np);  // prop.cu:702
    if (rc) return rc;
    rc = ensure_prev(c, np);
    if (rc) return rc;
    ProFrame& F = bh->h_pro[i];
    F.W = c->W;
    F.H = c->H;
    F.npix = c->npix;
    F.np = np;
    F.nbuf = 10 * np;
    F.sp_size = c->cfg.sp_size;
    F.flow = flows[i];
    F.spix_prev = spix_prev[i];
    F.mu_shape_prev = prevs[i].mu_shape;
    F.prior_counts = prevs[i].prior_counts;
    F.invalid_counts = prevs[i].invalid_counts;
    memcpy(&F.in, &prevs[i], sizeof(F.in));
    F.P = c->P;
    F.fsum = c->fsum;
    F.counts_prev = c->counts_prev;
    F.counts_s = c->counts_s;
    F.counts_f = c->counts_f;
    F.flow_ds = c->flow_ds;
    F.centers = c->centers;
    F.keys = c->keys;
    F.shifted = c->shifted;
    F.seg = c->seg;
    F.parent = c->parent;
    F.rsize = c->rsize;
    F.max_spix = c->max_spix;
    F.border = c->border;
    F.best = c->best;
    for (int k = 0; k < kProKinds; k++)
      off[k * (kProMax + 1) + i + 1] =
          off[k * (kProMax + 1) + i] + pro_blocks(k, c->W, c->H, np, 10 * np);
  }
  int total[kProKinds];
  for (int k = 0; k < kProKinds; k++) total[k] = off[k * (kProMax + 1) + n];
  HIP_TRY(hipMemcpyAsync(bh->d_pro, bh->h_pro, sizeof(ProFrame) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(bh->d_pro_off, off, sizeof(int) * kProKinds * (kProMax + 1),
                         hipMemcpyHostToDevice, s));
  Pro p;
  p.fr = bh->d_pro;
  p.off = bh->d_pro_off;
  p.nf = n;
  p.nall = n;
  p.fillres = bh->d_fillres;
  p.negs = bh->d_negs;
  p.hlen = bh->d_hlen;
  pro_shift(p, total, s);
  int rc = fill_loop_pro(bh, p, total, s);
  if (rc) return rc;
  pro_disconnected(p, total, s);
  rc = fill_loop_pro(bh, p, total, s);
  if (rc) return rc;
  pro_finish(p, total, s);
  for (int i = 0; i < n; i++) {
    const size_t bytes = sizeof(int) * (size_t)cs[i]->npix;
    if (shifted_out && shifted_out[i])
      HIP_TRY(hipMemcpyAsync(shifted_out[i], cs[i]->shifted, bytes, hipMemcpyDeviceToDevice, s));
    if (filled_out && filled_out[i])
      HIP_TRY(hipMemcpyAsync(filled_out[i], cs[i]->seg, bytes, hipMemcpyDeviceToDevice, s));
  }
  HIP_TRY(hipMemcpyAsync(bh->h_negs, p.negs, sizeof(int) * n * kNegSub, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = cs[i];
    const stsp_config& cfg = c->cfg;
    const int np = nprev[i];
    // adaptive iteration count (pyapi.cu:139-149): holes of the shifted map
    long long holes = 0;
    for (int j = 0; j < kNegSub; j++) holes += bh->h_negs[i * kNegSub + j];
    float iperc = (float)(holes / (1.0 * c->npix));
    int adapt = iperc > 0.20f ? 12 : (iperc < 0.01f ? 4 : 8);
    if (niters_state) niters_state[i] = adapt;
    c->img = imgs[i];
    c->nbuf = 10 * np;
    c->nspix_prev = np;
    c->prop_mode = 1;
    c->split_alpha = cfg.split_alpha;
    c->merge_alpha = cfg.merge_alpha;
    c->thresh_new = cfg.thresh_new;
    c->niters = cfg.prop_niters > 0 ? cfg.prop_niters : adapt;
    c->prev_prior_mu_app = prevs[i].prior_mu_app;
  }
  HIP_TRY(hipGetLastError());
  return STSP_OK;
}

// compaction (compact_spix.cu:41-122 as called at prop.cu:820-829 and
// bass.cu:319-320); the new-label count is copied to pinned memory
static int epilogue_compact(stsp_ctx* c, hipStream_t s) {
  launch_compact(c->seg, c->npix, &c->P, c->nbuf, c->nspix_prev, c->nbuf, c->alive, c->flags,
                 c->rank, c->num_new, c->ctmp, s);
  HIP_TRY(hipMemcpyAsync(c->host + 1, c->num_new, sizeof(int), hipMemcpyDeviceToHost, s));
  return STSP_OK;
}

// export after the stream has been synchronised (sparams_io.cu:65-79, 583-688)
static int epilogue_finish(stsp_ctx* c, int* spix_out, stsp_params* out, int* nspix_out,
                           hipStream_t s) {
  int nspix = c->nspix_prev + c->host[1];
  *nspix_out = nspix;
  if (nspix > out->cap) return STSP_ERR_CAPACITY;
  k_set_int<<<1, 1, 0, s>>>(c->total, nspix);
  launch_fill(c->alive, 0, c->nbuf, s);
  launch_mark_alive(c->seg, c->npix, c->alive, c->nbuf, s);
  OutParams o;
  memcpy(&o, out, sizeof(o));
  launch_export(c->P, o, c->total, c->alive, nspix, s);
  HIP_TRY(hipMemcpyAsync(spix_out, c->seg, sizeof(int) * c->npix, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipGetLastError());
  return STSP_OK;
}

static int run_single(stsp_ctx* c, int* spix_out, stsp_params* out, int* nspix_out,
                      hipStream_t s) {
  stsp_ctx* one = c;
  int rc = refine(&c->self, &one, 1, s);
  if (!rc) rc = epilogue_compact(c, s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  rc = epilogue_finish(c, spix_out, out, nspix_out, s);
  if (!rc) rc = mark_done(c, s);
  if (rc) return rc;
  if (c->timing) {
    HIP_TRY(hipStreamSynchronize(s));
    tflush(c);
  }
  return STSP_OK;
}

// ---------------------------------------------------------------------------
// single-image BASS without allocation (stsp_bass_noalloc)
// ---------------------------------------------------------------------------
// Every device that runs the entry point holds its own copy of `scratch` for
// the life of the process.  In host code `scratch` is only a stand-in for
// the symbol: the device address of each copy comes from
// hipGetSymbolAddress, taken once per device (with that device current) on
// its first call, and only that address is handed to kernels and copies.
// All calls on one device share the buffer and keep the last call's
// contents; calls that overlap on one device race (a host mutex serialises
// the host side, so the device work of two calls must also be ordered, e.g.
// on one stream).
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];

struct NoAllocDevice {
  std::mutex mu;
  char* ptr = nullptr;  // this device's copy of `scratch`
  stsp_ctx ctx;
  int host[16];                         // read-back values (pageable)
  FrameDesc h_fr[1];                    // batch table of one frame
  int h_off[kOffArrays * 2];
};
static NoAllocDevice g_noalloc[kScratchDevices];

// BASS capacity rule of bass_prologue_n (30 K0 slots) plus the growth
// headroom of ensure_labels
static int noalloc_cap(const stsp_config& cfg, int H, int W) {
  HexGeom g = hex_geometry(cfg.sp_size, W, H);
  const int nbuf = 30 * g.nx * g.ny;
  return nbuf + nbuf / 4;
}

static size_t noalloc_tables(int cap) {
  return 256 + ((sizeof(FrameDesc) + 255) & ~(size_t)255) + 256 + 256 +
         ((sizeof(int) * (size_t)kWorkLists * cap + 255) & ~(size_t)255);
}

static size_t noalloc_bytes(const stsp_config& cfg, int H, int W) {
  const int cap = noalloc_cap(cfg, H, W);
  return ((pixel_bytes((size_t)H * W) + 255) & ~(size_t)255) + label_bytes(cap) +
         noalloc_tables(cap);
}

static int noalloc_setup(NoAllocDevice& d) {
  if (d.ptr) return STSP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return STSP_ERR_HIP;
  d.ptr = (char*)p;
  return STSP_OK;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Lays the context and its batch of one out over the device's scratch copy.
static stsp_ctx* noalloc_context(NoAllocDevice& d, const stsp_config& cfg, int H, int W,
                                 hipStream_t s) {
  stsp_ctx* c = &d.ctx;
  c->cfg = cfg;
  c->H = H;
  c->W = W;
  c->npix = H * W;
  c->lut = make_lut();
  c->noalloc = true;
  c->timing = 0;
  c->sweep_shape = 4;
  c->wave_box = kWaveBoxMax;
  c->win_h = c->win_w = 0;
  c->shape_marginal = 0;
  c->nprev = 0;
  c->host = d.host;
  char* base = d.ptr;
  carve_pixels(c, base);
  base += (pixel_bytes((size_t)H * W) + 255) & ~(size_t)255;
  const int cap = noalloc_cap(cfg, H, W);
  (void)carve_labels(c, base, cap);  // sized by label_bytes: cannot overflow
  base += label_bytes(cap);
  stsp_batch* bh = &c->self;
  bh->noalloc = true;
  bh->use_graph = false;
  bh->maxf = 1;
  Arena a;
  a.base = base;
  a.size = noalloc_tables(cap);
  bh->d_work_n = a.take<int>(kWorkCounters);
  bh->d_fr = a.take<FrameDesc>(1);
  bh->d_off = a.take<int>(kOffArrays * 2);
  bh->d_work = a.take<int>((size_t)kWorkLists * cap);
  bh->work_cap = cap;
  bh->h_fr = d.h_fr;
  bh->h_off = d.h_off;
  k_zero_u64<<<1, 64, 0, s>>>(c->search_count, kSearchPaths);
  return c;
}

extern "C" {

void stsp_default_config(stsp_config* cfg) {
  cfg->niters = 500;
  cfg->niters_seg = 4;
  cfg->sm_start = 0;
  cfg->sp_size = 25;
  cfg->sm_period = 4;
  cfg->prop_niters = 0;
  cfg->sigma2_app = 8.1e-5f;
  cfg->potts = 0.5f;
  cfg->alpha_hastings = logf(0.5f);
  cfg->split_alpha = 0.0f;
  cfg->merge_alpha = 0.0f;
  cfg->thresh_relabel = 1e-5f;
  cfg->thresh_new = 1e-2f;
  cfg->target_nspix = 0;
}

const char* stsp_status_string(int status) {
  switch (status) {
    case STSP_OK: return "ok";
    case STSP_ERR_ARG: return "invalid argument";
    case STSP_ERR_CAPACITY: return "output capacity too small";
    case STSP_ERR_FILL: return "hole filling stopped shrinking";
    case STSP_ERR_HIP: return "HIP runtime error";
    case STSP_ERR_STATE: return "invalid stream state";
    case STSP_ERR_TOO_LARGE: return "frame exceeds the static scratch buffer";
    default: return "unknown status";
  }
}

int stsp_bass_capacity(int sp_size, int height, int width) {
  HexGeom g = hex_geometry(sp_size, width, height);
  return 30 * g.nx * g.ny;
}

int stsp_prop_capacity(int nspix_prev) { return 10 * nspix_prev; }

int stsp_ctx_create(const stsp_config* cfg, int height, int width, stsp_ctx** out) {
  if (!cfg || !out || height < 2 || width < 2 || cfg->sp_size < 1) return STSP_ERR_ARG;
  if ((long long)height * width >= (1ll << 23)) return STSP_ERR_ARG;
  stsp_ctx* c = new (std::nothrow) stsp_ctx();
  if (!c) return STSP_ERR_HIP;
  c->cfg = *cfg;
  c->H = height;
  c->W = width;
  c->npix = height * width;
  c->lut = make_lut();
  int rc = alloc_pixels(c);
  if (!rc) rc = batch_alloc(&c->self, 1);
  if (!rc && hipEventCreateWithFlags(&c->done_ev, hipEventDisableTiming) != hipSuccess)
    rc = STSP_ERR_HIP;
  if (!rc && hipMemset(c->search_count, 0, sizeof(unsigned long long) * kSearchPaths) != hipSuccess)
    rc = STSP_ERR_HIP;
  if (rc) {
    stsp_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return STSP_OK;
}

void stsp_ctx_destroy(stsp_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  for (int k = 0; k < STSP_KCLASS_COUNT; k++)
    for (int i = 0; i < 2 * c->timers[k].created; i++) (void)hipEventDestroy(c->timers[k].ev[i]);
  batch_free(&c->self);
  if (c->done_ev) (void)hipEventDestroy(c->done_ev);
  if (c->pix_mem) (void)hipFree(c->pix_mem);
  if (c->lab_mem) (void)hipFree(c->lab_mem);
  if (c->prev_mem) (void)hipFree(c->prev_mem);
  if (c->host) (void)hipHostFree(c->host);
  delete c;
}

int stsp_ctx_set_config(stsp_ctx* c, const stsp_config* cfg) {
  if (!c || !cfg || cfg->sp_size < 1) return STSP_ERR_ARG;
  c->cfg = *cfg;
  return STSP_OK;
}

int stsp_ctx_set_tuning(stsp_ctx* c, int sweep_shape) {
  if (!c || sweep_shape < 0 || sweep_shape >= kSweepShapes) return STSP_ERR_ARG;
  c->sweep_shape = sweep_shape;
  return STSP_OK;
}

int stsp_ctx_set_split_search(stsp_ctx* c, int wave_box, int win_rows, int win_cols) {
  if (!c || wave_box < 1 || wave_box > kWaveBoxMax || win_rows < 1 || win_rows > kSplitWinRows ||
      win_cols < 1 || win_cols > kSplitWinCols)
    return STSP_ERR_ARG;
  c->wave_box = wave_box;
  c->win_h = win_rows;
  c->win_w = win_cols;
  return STSP_OK;
}

int stsp_ctx_set_shape_marginal(stsp_ctx* c, int on) {
  if (!c || (on != 0 && on != 1)) return STSP_ERR_ARG;
  c->shape_marginal = on;
  return STSP_OK;
}

int stsp_ctx_device_bytes(stsp_ctx* c, size_t* bytes) {
  if (!c || !bytes) return STSP_ERR_ARG;
  *bytes = c->pix_bytes + c->lab_bytes + c->prev_bytes;
  return STSP_OK;
}

int stsp_ctx_split_search_counts(stsp_ctx* c, long long* counts) {
  if (!c || !counts) return STSP_ERR_ARG;
  unsigned long long h[kSearchPaths];
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(h, c->search_count, sizeof(h), hipMemcpyDeviceToHost));
  for (int i = 0; i < kSearchPaths; i++) counts[i] = (long long)h[i];
  return STSP_OK;
}

int stsp_ctx_enable_timing(stsp_ctx* c, int on) {
  if (!c) return STSP_ERR_ARG;
  c->timing = on ? 1 : 0;
  for (int k = 0; k < STSP_KCLASS_COUNT; k++) {
    c->timers[k].ms = 0.0;
    c->timers[k].count = 0;
    c->timers[k].n = 0;
  }
  return STSP_OK;
}

int stsp_ctx_kernel_time(stsp_ctx* c, int kclass, double* total_ms, long long* launches) {
  if (!c || kclass < 0 || kclass >= STSP_KCLASS_COUNT || !total_ms || !launches)
    return STSP_ERR_ARG;
  *total_ms = c->timers[kclass].ms;
  *launches = c->timers[kclass].count;
  return STSP_OK;
}

int stsp_bass(stsp_ctx* c, const float* img, int* spix_out, stsp_params* out, int* nspix_out,
              hipStream_t s) {
  if (!c || !img || !spix_out || !out || !nspix_out) return STSP_ERR_ARG;
  const float* im = img;
  int rc = bass_prologue_n(&c, &im, 1, nullptr, s);
  if (rc) return rc;
  return run_single(c, spix_out, out, nspix_out, s);
}

int stsp_prop_ex(stsp_ctx* c, const float* img, const float* flow, const int* spix_prev,
                 stsp_params* prev, int nspix_prev, int* spix_out, stsp_params* out,
                 int* nspix_out, int* niters_state, int* shifted_out, int* filled_out,
                 hipStream_t s) {
  if (!c || !img || !flow || !spix_prev || !prev || !spix_out || !out || !nspix_out ||
      nspix_prev < 1)
    return STSP_ERR_ARG;
  int rc = prop_prologue_n(&c->self, &c, 1, &img, &flow, &spix_prev, prev, &nspix_prev,
                           niters_state, s, &shifted_out, &filled_out);
  if (rc) return rc;
  return run_single(c, spix_out, out, nspix_out, s);
}

int stsp_prop(stsp_ctx* c, const float* img, const float* flow, const int* spix_prev,
              stsp_params* prev, int nspix_prev, int* spix_out, stsp_params* out,
              int* nspix_out, int* niters_state, hipStream_t s) {
  return stsp_prop_ex(c, img, flow, spix_prev, prev, nspix_prev, spix_out, out, nspix_out,
                      niters_state, nullptr, nullptr, s);
}

int stsp_batch_create(int max_frames, stsp_batch** out) {
  if (!out || max_frames < 1 || max_frames > 64) return STSP_ERR_ARG;
  stsp_batch* bh = new (std::nothrow) stsp_batch();
  if (!bh) return STSP_ERR_HIP;
  int rc = batch_alloc(bh, max_frames);
  if (rc) {
    stsp_batch_destroy(bh);
    return rc;
  }
  *out = bh;
  return STSP_OK;
}

void stsp_batch_destroy(stsp_batch* bh) {
  if (!bh) return;
  (void)hipDeviceSynchronize();
  batch_free(bh);
  delete bh;
}

int stsp_batch_time_sweeps(stsp_batch* bh, int nlaunch, hipStream_t s, float* ms_out) {
  if (!bh || !ms_out || nlaunch < 1 || bh->last_shape < 0) return STSP_ERR_ARG;
  hipEvent_t e[2];
  HIP_TRY(hipEventCreateWithFlags(&e[0], hipEventDisableSystemFence));
  HIP_TRY(hipEventCreateWithFlags(&e[1], hipEventDisableSystemFence));
  // compaction renumbered the labels after the last refinement: recount the
  // statistics and sweep records so the timed sweeps see a consistent state
  b_update_params(bh->b, 0, false, s);
  const int per = kSweepShape[bh->last_shape].s;
  HIP_TRY(hipEventRecord(e[0], s));
  for (int i = 0; i < nlaunch; i++) (void)b_sweeps(bh->b, 0, per, bh->last_shape, bh->last_lut, s);
  HIP_TRY(hipEventRecord(e[1], s));
  HIP_TRY(hipEventSynchronize(e[1]));
  const hipError_t er = hipEventElapsedTime(ms_out, e[0], e[1]);
  (void)hipEventDestroy(e[0]);
  (void)hipEventDestroy(e[1]);
  HIP_TRY(hipGetLastError());
  return er == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

int stsp_batch_tune_sweeps(stsp_batch* bh, int nlaunch, hipStream_t s, float* us_per_sweep,
                           int* best) {
  if (!bh || !us_per_sweep || !best || nlaunch < 1 || bh->last_shape < 0) return STSP_ERR_ARG;
  hipEvent_t e[2];
  HIP_TRY(hipEventCreateWithFlags(&e[0], hipEventDisableSystemFence));
  HIP_TRY(hipEventCreateWithFlags(&e[1], hipEventDisableSystemFence));
  float launch_us[kSweepShapes];
  hipError_t er = hipSuccess;
  for (int sh = 0; sh < kSweepShapes && er == hipSuccess; sh++) {
    // every shape starts from the same recount of the final labels
    b_update_params(bh->b, 0, false, s);
    const int per = kSweepShape[sh].s;
    (void)b_sweeps(bh->b, 0, per, sh, bh->last_lut, s);  // untimed first launch
    er = hipEventRecord(e[0], s);
    for (int i = 0; i < nlaunch && er == hipSuccess; i++)
      (void)b_sweeps(bh->b, 0, per, sh, bh->last_lut, s);
    if (er == hipSuccess) er = hipEventRecord(e[1], s);
    if (er == hipSuccess) er = hipEventSynchronize(e[1]);
    float ms = 0.f;
    if (er == hipSuccess) er = hipEventElapsedTime(&ms, e[0], e[1]);
    launch_us[sh] = ms * 1000.f / nlaunch;
    us_per_sweep[sh] = launch_us[sh] / per;
  }
  (void)hipEventDestroy(e[0]);
  (void)hipEventDestroy(e[1]);
  HIP_TRY(er);
  HIP_TRY(hipGetLastError());
  // the cost of one iteration's sweeps as b_sweeps runs them: launches of
  // the shape, the remainder with the one-sweep shape
  const int nseg = bh->last_nseg > 0 ? bh->last_nseg : 1;
  int arg = 0;
  float cost_best = 0.f;
  for (int sh = 0; sh < kSweepShapes; sh++) {
    const int per = kSweepShape[sh].s;
    const float cost = (nseg / per) * launch_us[sh] + (nseg % per) * launch_us[0];
    if (sh == 0 || cost < cost_best) {
      cost_best = cost;
      arg = sh;
    }
  }
  *best = arg;
  return STSP_OK;
}

int stsp_batch_use_graph(stsp_batch* bh, int on) {
  if (!bh) return STSP_ERR_ARG;
  bh->use_graph = on != 0;
  if (!on) invalidate_graph(bh);
  return STSP_OK;
}

}  // extern "C"

// Compaction and export of every frame of a batch (epilogue_compact and
// epilogue_finish_async for all frames, one launch per stage): one table
// upload, the launches, and one copy of the new-label counts, which the
// caller reads after its wait.
static int epilogue_batch(stsp_batch* bh, stsp_ctx** ctxs, int n, int** spix_out,
                          const stsp_params* outs, hipStream_t s) {
  int* off_lab = bh->h_epi_off;
  int* off_pix = bh->h_epi_off + (bh->maxf + 1);
  int lb = 0, pb = 0;
  for (int i = 0; i < n; i++) {
    stsp_ctx* c = ctxs[i];
    EpiFrame& E = bh->h_epi[i];
    E.P = c->P;
    memcpy(&E.o, &outs[i], sizeof(E.o));
    E.seg = c->seg;
    E.spix_out = spix_out[i];
    E.alive = c->alive;
    E.flags = c->flags;
    E.rank = c->rank;
    E.num_new = c->num_new;
    E.total = c->total;
    E.ctmp = c->ctmp;
    E.npix = c->npix;
    E.nbuf = c->nbuf;
    E.prev = c->nspix_prev;
    off_lab[i] = lb;
    off_pix[i] = pb;
    lb += div_up(c->nbuf, kBlock);
    pb += div_up(c->npix, kBlock);
  }
  off_lab[n] = lb;
  off_pix[n] = pb;
  HIP_TRY(hipMemcpyAsync(bh->d_epi, bh->h_epi, sizeof(EpiFrame) * n, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(bh->d_epi_off, bh->h_epi_off, sizeof(int) * 2 * (bh->maxf + 1),
                         hipMemcpyHostToDevice, s));
  EpiBatch e;
  e.fr = bh->d_epi;
  e.off_lab = bh->d_epi_off;
  e.off_pix = bh->d_epi_off + (bh->maxf + 1);
  e.counts = bh->d_epi_counts;
  e.nf = n;
  e.lab_blocks = lb;
  e.pix_blocks = pb;
  epi_launch(e, s);
  HIP_TRY(hipMemcpyAsync(bh->h_epi_counts, bh->d_epi_counts, sizeof(int) * n,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipGetLastError());
  return STSP_OK;
}

// Refines frames whose prologues are done: frames with the same schedule
// (iteration count, mode, sweeps per iteration, split/merge period and start)
// share one launch sequence; then compaction, one wait, and the exports.
static int refine_and_finish(stsp_batch* bh, stsp_ctx** ctxs, int n, int** spix_out,
                             stsp_params* outs, int* nspix_out, hipStream_t s) {
  std::vector<int> done(n, 0);
  std::vector<stsp_ctx*> group;
  // the work lists of the largest group are allocated before any group runs
  int work = 0;
  for (int i = 0; i < n; i++) work += ctxs[i]->cap;
  int rc0 = ensure_work(bh, work, s);
  if (rc0) return rc0;
  int base = 0, slot = 0;
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    group.clear();
    const stsp_ctx* a = ctxs[i];
    for (int j = i; j < n; j++) {
      const stsp_ctx* c = ctxs[j];
      if (done[j] || c->niters != a->niters || c->prop_mode != a->prop_mode ||
          c->cfg.niters_seg != a->cfg.niters_seg || c->cfg.sm_period != a->cfg.sm_period ||
          c->cfg.sm_start != a->cfg.sm_start)
        continue;
      group.push_back(ctxs[j]);
      done[j] = 1;
    }
    int rc = refine(bh, group.data(), (int)group.size(), s, base, slot);
    if (rc) return rc;
    base += (int)group.size();
    slot++;
  }
  int rc = epilogue_batch(bh, ctxs, n, spix_out, outs, s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));  // the counts (and the refinement's timing events)
  for (int i = 0; i < n; i++) tflush(ctxs[i]);  // timed launches are recorded on group leads
  for (int i = 0; i < n; i++) {
    rc = mark_done(ctxs[i], s);
    if (rc) return rc;
  }
  for (int i = 0; i < n; i++) {
    ctxs[i]->host[1] = bh->h_epi_counts[i];
    nspix_out[i] = ctxs[i]->nspix_prev + ctxs[i]->host[1];
    if (nspix_out[i] > outs[i].cap) return STSP_ERR_CAPACITY;
  }
  return STSP_OK;
}

static int check_batch_ctxs(stsp_batch* bh, stsp_ctx** ctxs, int n) {
  if (!bh || !ctxs || n < 1 || n > bh->maxf) return STSP_ERR_ARG;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || ctxs[i]->cfg.target_nspix > 0) return STSP_ERR_ARG;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return STSP_ERR_ARG;
  }
  return STSP_OK;
}

extern "C" {

int stsp_prop_batch(stsp_batch* bh, stsp_ctx** ctxs, int n, const float** imgs,
                    const float** flows, const int** spix_prev, stsp_params* prevs,
                    const int* nspix_prev, int** spix_out, stsp_params* outs, int* nspix_out,
                    int* niters_state, hipStream_t s) {
  int rc = check_batch_ctxs(bh, ctxs, n);
  if (rc) return rc;
  if (!imgs || !flows || !spix_prev || !prevs || !nspix_prev || !spix_out || !outs || !nspix_out)
    return STSP_ERR_ARG;
  for (int i = 0; i < n; i++)
    if (nspix_prev[i] < 1) return STSP_ERR_ARG;
  rc = prop_prologue_n(bh, ctxs, n, imgs, flows, spix_prev, prevs, nspix_prev, niters_state, s);
  if (rc) return rc;
  return refine_and_finish(bh, ctxs, n, spix_out, outs, nspix_out, s);
}

int stsp_bass_batch(stsp_batch* bh, stsp_ctx** ctxs, int n, const float** imgs, int** spix_out,
                    stsp_params* outs, int* nspix_out, hipStream_t s) {
  int rc = check_batch_ctxs(bh, ctxs, n);
  if (rc) return rc;
  if (!imgs || !spix_out || !outs || !nspix_out) return STSP_ERR_ARG;
  rc = bass_prologue_n(ctxs, imgs, n, &bh->lanes, s);
  if (rc) return rc;
  return refine_and_finish(bh, ctxs, n, spix_out, outs, nspix_out, s);
}

int stsp_rescale_image(const uint8_t* img_u8, float* out, int nimages, int npix, float scale,
                       hipStream_t stream) {
  if (!img_u8 || !out || nimages < 1 || npix < 1 || !(scale != 0.0f)) return STSP_ERR_ARG;
  launch_rescale_u8(img_u8, out, 3ll * nimages * npix, scale, stream);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

int stsp_rgb2lab(const float* img_bgr, float* img_lab, int nimages, int npix, hipStream_t stream) {
  if (!img_bgr || !img_lab || nimages < 1 || npix < 1) return STSP_ERR_ARG;
  launch_rgb_to_lab(img_bgr, img_lab, (long long)nimages * npix, stream);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

int stsp_lab2rgb(const float* img_lab, float* img_rgb, int nimages, int npix, hipStream_t stream) {
  if (!img_lab || !img_rgb || nimages < 1 || npix < 1) return STSP_ERR_ARG;
  launch_lab_to_rgb(img_lab, img_rgb, (long long)nimages * npix, stream);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

int stsp_bass_noalloc_scratch_bytes(const stsp_config* cfg, int height, int width,
                                    size_t* bytes) {
  if (!cfg || !bytes || height < 2 || width < 2 || cfg->sp_size < 1) return STSP_ERR_ARG;
  *bytes = noalloc_bytes(*cfg, height, width);
  return STSP_OK;
}

int stsp_bass_noalloc(const stsp_config* cfg, int height, int width, const float* img,
                      int* spix_out, stsp_params* out, int* nspix_out, hipStream_t s) {
  if (!cfg || !img || !spix_out || !out || !nspix_out || height < 2 || width < 2 ||
      cfg->sp_size < 1 || cfg->target_nspix > 0)
    return STSP_ERR_ARG;
  if ((long long)height * width >= (1ll << 23)) return STSP_ERR_ARG;
  if (noalloc_bytes(*cfg, height, width) > kScratchBytes) return STSP_ERR_TOO_LARGE;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return STSP_ERR_ARG;
  NoAllocDevice& d = g_noalloc[dev];
  std::lock_guard<std::mutex> lock(d.mu);
  int rc = noalloc_setup(d);
  if (rc) return rc;
  stsp_ctx* c = noalloc_context(d, *cfg, height, width, s);
  const float* im = img;
  rc = bass_prologue_n(&c, &im, 1, nullptr, s);
  if (rc) return rc;
  return run_single(c, spix_out, out, nspix_out, s);
}

int stsp_marked_video(const float* vid, const int* spix, const float* color_host, int nframes,
                      int height, int width, float* marked, hipStream_t stream) {
  if (!vid || !spix || !color_host || !marked || nframes < 1 || height < 1 || width < 1)
    return STSP_ERR_ARG;
  launch_marked(vid, spix, color_host, nframes, height, width, marked, stream);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

}  // extern "C"
