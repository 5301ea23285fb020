// Host orchestration of the BASS/BIST frame loop and the C ABI (stspix.h).
//
// One stsp_ctx owns the device workspace for a frame size.  A frame runs as
// a sequence of kernel launches on the caller's stream; the only host waits
// are the ones that size later launches (label counts after the first
// compaction, the hole-filling loop and the final superpixel count).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "../../include/stspix.h"
#include "kernels.h"

using namespace stsp;

#define HIP_TRY(x)                                   \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) return STSP_ERR_HIP;       \
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

}  // namespace

// Event-pair timer for one kernel class (measures device time of the
// launches it brackets on the launch stream).
struct KTimer {
  static const int kMax = 1024;
  hipEvent_t ev[2 * kMax];
  int created = 0, n = 0;
  double ms = 0.0;
  long long count = 0;
};

struct stsp_ctx {
  stsp_config cfg;
  int timing = 0;
  KTimer timers[STSP_KCLASS_COUNT];
  int H, W, npix;
  int cap = 0;    // label slots of the working parameter set
  int nprev = 0;  // label slots of the previous-frame arrays
  SimpleLut lut;
  // pixel buffers
  int *seg, *seg2, *halves, *d1, *d2, *list, *shifted, *parent, *rsize;
  int sweep_shape = 2;  // 0: 1 sweep/launch, 1: 2, 2: 4 (fused)
  uint8_t* border;
  unsigned long long* keys;
  // label buffers
  Params P;
  LabelStats st, sst;
  SmState sm;
  SweepRec* rec;
  int *off, *cursor, *alive, *flags, *rank, *target, *want_new, *src_of;
  void* ctmp;
  // previous-frame buffers
  int64_t* fsum;
  int *counts_prev, *counts_s, *counts_f;
  float* flow_ds;
  double* centers;
  unsigned long long* best;
  // scalars
  int *max_spix, *total, *num_new, *num_neg;
  int* host;  // pinned
  void* pix_mem = nullptr;
  void* lab_mem = nullptr;
  void* prev_mem = nullptr;
};

static int alloc_pixels(stsp_ctx* c) {
  size_t n = c->npix;
  Arena a;
  a.size = 13 * 256 + n * (4 * 9 + 1 + 8) + 4096;
  HIP_TRY(hipMalloc(&c->pix_mem, a.size));
  a.base = (char*)c->pix_mem;
  c->seg = a.take<int>(n);
  c->seg2 = a.take<int>(n);
  c->halves = a.take<int>(n);
  c->d1 = a.take<int>(n);
  c->d2 = a.take<int>(n);
  c->list = a.take<int>(n);
  c->shifted = a.take<int>(n);
  c->parent = a.take<int>(n);
  c->rsize = a.take<int>(n);
  c->border = a.take<uint8_t>(n);
  c->keys = a.take<unsigned long long>(n);
  c->max_spix = a.take<int>(1);
  c->total = a.take<int>(1);
  c->num_new = a.take<int>(1);
  c->num_neg = a.take<int>(1);
  HIP_TRY(hipHostMalloc((void**)&c->host, 16 * sizeof(int), hipHostMallocDefault));
  return STSP_OK;
}

// (re)allocate the per-label workspace for `cap` slots
static int ensure_labels(stsp_ctx* c, int cap) {
  if (cap <= c->cap) return STSP_OK;
  if (c->lab_mem) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(c->lab_mem));
    c->lab_mem = nullptr;
  }
  cap = cap + cap / 4;  // headroom for growth across frames
  size_t n = cap, nsm = 2 * n + 2;
  // byte budget, generously rounded (each take pads to 256 B)
  size_t bytes = n * (3 * 4 * 2 + 8 * (2 + 3 + 2 + 3 + 3 + 3 + 1) + 4 * 6) +  // params
                 n * (4 + 8 * 11 + 4 + 16) +                                   // st
                 nsm * (4 + 8 * 11 + 4) +                                      // sst
                 nsm * (4 * 4 + 8 * 3 * 2 + 8 * 2 + 4 * 8) + nsm * 8 +          // sm
                 n * sizeof(SweepRec) + n * 4 * 8 + compact_tmp_bytes(cap) + 96 * 256;
  HIP_TRY(hipMalloc(&c->lab_mem, bytes));
  Arena a;
  a.base = (char*)c->lab_mem;
  a.size = bytes;
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
  c->off = a.take<int>(n);
  c->cursor = a.take<int>(n);
  c->alive = a.take<int>(n);
  c->flags = a.take<int>(n);
  c->rank = a.take<int>(n);
  c->target = a.take<int>(n);
  c->want_new = a.take<int>(n);
  c->src_of = a.take<int>(n);
  c->ctmp = (void*)a.take<char>(compact_tmp_bytes(cap));
  if (a.used > a.size) {
    fprintf(stderr, "stspix: label arena overflow (%zu > %zu)\n", a.used, a.size);
    return STSP_ERR_HIP;
  }
  c->cap = cap;
  return STSP_OK;
}

static int ensure_prev(stsp_ctx* c, int n) {
  if (n <= c->nprev) return STSP_OK;
  if (c->prev_mem) {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipFree(c->prev_mem));
    c->prev_mem = nullptr;
  }
  n = n + n / 4;
  size_t bytes = (size_t)n * (16 + 12 + 8 + 16 + 8) + 16 * 256;
  HIP_TRY(hipMalloc(&c->prev_mem, bytes));
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

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static void tstart(stsp_ctx* c, int cls, hipStream_t s) {
  if (!c->timing) return;
  KTimer& t = c->timers[cls];
  if (t.n >= KTimer::kMax) return;
  if (t.created <= t.n) {
    (void)hipEventCreate(&t.ev[2 * t.n]);
    (void)hipEventCreate(&t.ev[2 * t.n + 1]);
    t.created = t.n + 1;
  }
  (void)hipEventRecord(t.ev[2 * t.n], s);
}
static void tstop(stsp_ctx* c, int cls, hipStream_t s) {
  if (!c->timing) return;
  KTimer& t = c->timers[cls];
  if (t.n >= KTimer::kMax) return;
  (void)hipEventRecord(t.ev[2 * t.n + 1], s);
  t.n++;
}
// called after the stream has been synchronised
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

static int read_int(stsp_ctx* c, const int* dev, hipStream_t s, int* out) {
  HIP_TRY(hipMemcpyAsync(c->host, dev, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *out = c->host[0];
  return STSP_OK;
}

static void update_params(stsp_ctx* c, const float* img, int nbuf, hipStream_t s,
                          bool bbox = false) {
  tstart(c, STSP_KCLASS_STATS, s);
  launch_stats(img, c->seg, nullptr, c->npix, c->W, c->st, nbuf, s, bbox);
  tstop(c, STSP_KCLASS_STATS, s);
  launch_param_update(c->P, c->st, c->rec, nbuf, c->cfg.sp_size, 0, s);
}

static void split_step(stsp_ctx* c, const float* img, const int* shifted, int nbuf,
                       int direction, const NigConsts& nc, int prop_mode, hipStream_t s) {
  int nsm = 2 * nbuf + 2;
  launch_sm_clear(c->sm, c->sst, nsm, nbuf, s);
  launch_split_proposal(c->seg, c->P, c->st.bbox, nbuf, c->max_spix, direction, c->W, c->H,
                        c->d1, c->d2, c->halves, s);
  launch_stats(img, c->halves, shifted, c->npix, c->W, c->sst, nsm, s);
  launch_split_eval_apply(c->seg, c->halves, c->P, c->st, c->sst, c->sm, nbuf, c->max_spix, nc,
                          c->cfg.split_alpha, prop_mode, c->npix, c->total, s);
}

// niters_seg boundary sweeps; the result lands in c->seg
static void sweeps(stsp_ctx* c, const float* img, hipStream_t s) {
  int* res = launch_sweeps(img, c->seg, c->seg2, c->rec, c->cfg.niters_seg, c->cfg.sigma2_app,
                           c->cfg.potts, c->H, c->W, c->lut, c->sweep_shape, s);
  if (res != c->seg) {
    c->seg2 = c->seg;
    c->seg = res;
  }
}

static inline bool is_split_iter(const stsp_config& cfg, int idx) {
  int p = cfg.sm_period > 0 ? cfg.sm_period : 4;
  return idx > cfg.sm_start && idx % p == 0;
}
static inline bool is_merge_iter(const stsp_config& cfg, int idx) {
  int p = cfg.sm_period > 0 ? cfg.sm_period : 4;
  return idx > cfg.sm_start && idx % p == p / 2;
}

static int finish_frame(stsp_ctx* c, int nbuf, int prev_nspix, int* spix_out, stsp_params* out,
                        int* nspix_out, hipStream_t s) {
  const float* dummy = nullptr;
  (void)dummy;
  launch_store_sample(c->P, c->st, nbuf, s);
  launch_compact(c->seg, c->npix, &c->P, nbuf, prev_nspix, nbuf, c->alive, c->flags, c->rank,
                 c->num_new, c->ctmp, s);
  int num_new = 0;
  int rc = read_int(c, c->num_new, s, &num_new);
  if (rc) return rc;
  int nspix = prev_nspix + num_new;
  *nspix_out = nspix;
  if (nspix > out->cap) return STSP_ERR_CAPACITY;
  k_set_int<<<1, 1, 0, s>>>(c->total, nspix);
  launch_fill(c->alive, 0, nbuf, s);
  launch_mark_alive(c->seg, c->npix, c->alive, nbuf, s);
  OutParams o;
  memcpy(&o, out, sizeof(o));
  launch_export(c->P, o, c->total, c->alive, nspix, s);
  HIP_TRY(hipMemcpyAsync(spix_out, c->seg, sizeof(int) * c->npix, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipGetLastError());
  if (c->timing) {
    HIP_TRY(hipStreamSynchronize(s));
    tflush(c);
  }
  return STSP_OK;
}

static int fill_loop(stsp_ctx* c, hipStream_t s) {
  int prev_neg = -1;
  for (int iter = 0; iter < 1000; iter++) {
    k_set_int<<<1, 1, 0, s>>>(c->num_neg, 0);
    launch_fill_count(c->seg, c->border, c->num_neg, c->W, c->H, s);
    int neg = 0;
    int rc = read_int(c, c->num_neg, s, &neg);
    if (rc) return rc;
    if (neg == 0) return STSP_OK;
    if (iter > 0 && neg == prev_neg) return STSP_ERR_FILL;
    launch_fill_phases(c->seg, c->border, c->centers, c->W, c->H, s);
    prev_neg = neg;
  }
  return STSP_OK;
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
}

const char* stsp_status_string(int status) {
  switch (status) {
    case STSP_OK: return "ok";
    case STSP_ERR_ARG: return "invalid argument";
    case STSP_ERR_CAPACITY: return "output capacity too small";
    case STSP_ERR_FILL: return "hole filling stopped shrinking";
    case STSP_ERR_HIP: return "HIP runtime error";
    case STSP_ERR_STATE: return "invalid stream state";
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
  if (c->pix_mem) (void)hipFree(c->pix_mem);
  if (c->lab_mem) (void)hipFree(c->lab_mem);
  if (c->prev_mem) (void)hipFree(c->prev_mem);
  if (c->host) (void)hipHostFree(c->host);
  delete c;
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

int stsp_ctx_set_tuning(stsp_ctx* c, int sweep_shape) {
  if (!c || sweep_shape < 0 || sweep_shape > 2) return STSP_ERR_ARG;
  c->sweep_shape = sweep_shape;
  return STSP_OK;
}

int stsp_ctx_set_config(stsp_ctx* c, const stsp_config* cfg) {
  if (!c || !cfg || cfg->sp_size < 1) return STSP_ERR_ARG;
  c->cfg = *cfg;
  return STSP_OK;
}

int stsp_bass(stsp_ctx* c, const float* img, int* spix_out, stsp_params* out, int* nspix_out,
              hipStream_t s) {
  if (!c || !img || !spix_out || !out || !nspix_out) return STSP_ERR_ARG;
  const stsp_config& cfg = c->cfg;
  int H = c->H, W = c->W, npix = c->npix;
  HexGeom g = hex_geometry(cfg.sp_size, W, H);
  int K0 = g.nx * g.ny;
  int nbuf = 30 * K0;
  int rc = ensure_labels(c, nbuf);
  if (rc) return rc;
  launch_hex_init(c->seg, H, W, cfg.sp_size, s);
  launch_compact(c->seg, npix, nullptr, K0, 0, nbuf, c->alive, c->flags, c->rank, c->num_new,
                 c->ctmp, s);
  int nspix = 0;
  rc = read_int(c, c->num_new, s, &nspix);
  if (rc) return rc;
  launch_bass_params_init(c->P, nspix, nbuf, cfg.sp_size, npix, c->st, img, c->seg, W, s);
  launch_border(c->seg, c->border, H, W, s);
  k_set_int<<<1, 1, 0, s>>>(c->max_spix, nspix - 1);
  NigConsts nc = make_nig(cfg);
  int smc = 1;
  for (int idx = 0; idx < cfg.niters; idx++) {
    update_params(c, img, nbuf, s, is_split_iter(cfg, idx));
    if (is_split_iter(cfg, idx)) {
      smc += 1;
      split_step(c, img, nullptr, nbuf, smc % 2 + 1, nc, 0, s);
      update_params(c, img, nbuf, s);
    }
    if (is_merge_iter(cfg, idx)) {
      launch_border(c->seg, c->border, H, W, s);
      launch_merge(c->seg, c->border, c->P, c->st, c->sst, c->sm, nbuf, 2 * nbuf + 2, nc, 0.0f,
                   smc % 2, 0, img, W, H, s);
      update_params(c, img, nbuf, s);
    }
    tstart(c, STSP_KCLASS_SWEEP, s);
    sweeps(c, img, s);
    tstop(c, STSP_KCLASS_SWEEP, s);
  }
  update_params(c, img, nbuf, s);
  return finish_frame(c, nbuf, 0, spix_out, out, nspix_out, s);
}

int stsp_prop(stsp_ctx* c, const float* img, const float* flow, const int* spix_prev,
              stsp_params* prev, int nspix_prev, int* spix_out, stsp_params* out,
              int* nspix_out, int* niters_state, hipStream_t s) {
  if (!c || !img || !flow || !spix_prev || !prev || !spix_out || !out || !nspix_out ||
      nspix_prev < 1)
    return STSP_ERR_ARG;
  const stsp_config& cfg = c->cfg;
  int H = c->H, W = c->W, npix = c->npix, n = nspix_prev;
  int nbuf = 10 * n;
  int rc = ensure_labels(c, nbuf);
  if (rc) return rc;
  rc = ensure_prev(c, n);
  if (rc) return rc;
  // shift_and_fill (shift_and_fill.cu:108-339)
  launch_flow_pool(flow, spix_prev, npix, c->fsum, c->counts_prev, prev->mu_shape, c->flow_ds,
                   c->centers, n, s);
  launch_shift(spix_prev, c->flow_ds, c->counts_prev, c->keys, c->shifted, W, H, s);
  HIP_TRY(hipMemcpyAsync(c->seg, c->shifted, sizeof(int) * npix, hipMemcpyDeviceToDevice, s));
  launch_count_labels(c->shifted, npix, c->counts_s, n, s);
  rc = fill_loop(c, s);
  if (rc) return rc;
  launch_invalidate_disconnected(c->seg, c->parent, c->rsize, c->best, W, H, n, s);
  rc = fill_loop(c, s);
  if (rc) return rc;
  launch_count_labels(c->seg, npix, c->counts_f, n, s);
  launch_prior_update(c->counts_prev, c->counts_s, c->counts_f, prev->prior_counts,
                      prev->invalid_counts, n, s);
  // adaptive iteration count (pyapi.cu:139-149)
  k_set_int<<<1, 1, 0, s>>>(c->num_neg, 0);
  launch_fill_count(c->shifted, c->border, c->num_neg, W, H, s);
  int ninvalid = 0;
  rc = read_int(c, c->num_neg, s, &ninvalid);
  if (rc) return rc;
  float iperc = (float)(ninvalid / (1.0 * npix));
  int adapt = iperc > 0.20f ? 12 : (iperc < 0.01f ? 4 : 8);
  if (niters_state) *niters_state = adapt;
  int niters = cfg.prop_niters > 0 ? cfg.prop_niters : adapt;
  // run_prop / prop_bass_v3 (prop.cu:326-468, 677-865)
  InParams in;
  memcpy(&in, prev, sizeof(in));
  launch_prop_params_init(c->P, in, n, nbuf, cfg.sp_size, c->seg, npix, s);
  k_set_int<<<1, 1, 0, s>>>(c->max_spix, n - 1);
  NigConsts nc = make_nig(cfg);
  int smc = 1;
  for (int idx = 0; idx < niters; idx++) {
    update_params(c, img, nbuf, s, is_split_iter(cfg, idx));
    if (is_split_iter(cfg, idx)) {
      smc += 1;
      split_step(c, img, c->shifted, nbuf, smc % 2 + 1, nc, 1, s);
      update_params(c, img, nbuf, s);
    }
    if (is_merge_iter(cfg, idx)) {
      launch_relabel(c->seg, npix, c->P, nbuf, prev->prior_mu_app, n, c->max_spix,
                     cfg.thresh_relabel, cfg.thresh_new, c->alive, c->target, c->want_new,
                     c->rank, c->src_of, c->total, s);
      smc += 1;
      launch_border(c->seg, c->border, H, W, s);
      launch_merge(c->seg, c->border, c->P, c->st, c->sst, c->sm, nbuf, 2 * nbuf + 2, nc,
                   cfg.merge_alpha, smc % 2, 1, img, W, H, s);
      update_params(c, img, nbuf, s);
    }
    tstart(c, STSP_KCLASS_SWEEP, s);
    sweeps(c, img, s);
    tstop(c, STSP_KCLASS_SWEEP, s);
  }
  update_params(c, img, nbuf, s);
  return finish_frame(c, nbuf, n, spix_out, out, nspix_out, s);
}

}  // extern "C"

extern "C" int stsp_marked_video(const float* vid, const int* spix, const float* color_host,
                                 int nframes, int height, int width, float* marked,
                                 hipStream_t stream) {
  if (!vid || !spix || !color_host || !marked || nframes < 1 || height < 1 || width < 1)
    return STSP_ERR_ARG;
  launch_marked(vid, spix, color_host, nframes, height, width, marked, stream);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}
