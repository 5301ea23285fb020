// Launch wrappers of the BASS/BIST kernels (internal to the library).
#pragma once

#include "common.h"

namespace stsp {

struct HexGeom {
  double Hh, w;
  int nx, ny;
};
HexGeom hex_geometry(int sp_size, int width, int height);

// 256-entry simple-point table packed in bits (bit j of word i = entry 32i+j)
struct SimpleLut {
  uint32_t bits[8];
};

// caller-visible parameter arrays (the reference SuperpixelParams field set)
struct OutParams {
  float* mu_app;
  float* prior_mu_app;
  double* mu_shape;
  double* sigma_shape;
  float* logdet_sigma_shape;
  double* prior_mu_shape;
  double* prior_sigma_shape;
  double* sample_sigma_shape;
  int* counts;
  int* sm_counts;
  float* prior_counts;
  int* ids;
  float* invalid_counts;
  int cap;
};
typedef OutParams InParams;

// k_bass.hip
void launch_hex_init(int* seg, int H, int W, int sp_size, hipStream_t s);
void launch_border(const int* seg, uint8_t* border, int H, int W, hipStream_t s);
void launch_update_seg(const float* img, int* seg, uint8_t* border, const SweepRec* rec,
                       int niters_seg, float sigma2_app, float potts, int H, int W,
                       const SimpleLut& lut, hipStream_t s);
void launch_stats(const float* img, const int* seg, const int* shifted, int npix, int W,
                  LabelStats st, int nlab, hipStream_t s, bool bbox = false);
void launch_param_update(Params P, LabelStats st, SweepRec* rec, int nbuf, int sp_size,
                         int summary_mode, hipStream_t s);
void launch_store_sample(Params P, LabelStats st, int nbuf, hipStream_t s);
void launch_bass_params_init(Params P, int nspix, int nbuf, int sp_size, int npix,
                             LabelStats st, const float* img, const int* seg, int W,
                             hipStream_t s);
void launch_scan(const int* in, int* out, int n, int* total, hipStream_t s);
void launch_fill(int* p, int v, int n, hipStream_t s);
void launch_mark_alive(const int* seg, int npix, int* alive, int n, hipStream_t s);
void launch_compact(int* seg, int npix, Params* P, int n, int prev, int nbuf, int* alive,
                    int* flags, int* rank, int* num_new, void* tmp, hipStream_t s);
size_t compact_tmp_bytes(int n);
void launch_export(Params P, OutParams o, const int* nspix, const int* alive, int nmax,
                   hipStream_t s);

// k_sm.hip
void launch_sm_clear(SmState sm, LabelStats sst, int nsm, int nbuf, hipStream_t s);
void launch_split_proposal(const int* seg, Params P, const int* bbox, int nbuf,
                           const int* max_spix, int direction, int W, int H, int* d1, int* d2,
                           int* halves, hipStream_t s);
void launch_split_eval_apply(int* seg, const int* halves, Params P, LabelStats st,
                             LabelStats sst, SmState sm, int nbuf, int* max_spix,
                             const NigConsts& nc, float split_alpha, int prop_mode, int npix,
                             int* total, hipStream_t s);
void launch_merge(int* seg, const uint8_t* border, Params P, LabelStats st, LabelStats sst,
                  SmState sm, int nbuf, int nsm, const NigConsts& nc, float merge_alpha,
                  int direction, int prop_mode, const float* img, int W, int H, hipStream_t s);

// k_prop.hip
void launch_flow_pool(const float* flow, const int* spix, int npix, int64_t* fsum, int* counts,
                      const double* mu_shape, float* flow_ds, double* centers, int n,
                      hipStream_t s);
void launch_shift(const int* spix, const float* flow_ds, const int* sizes,
                  unsigned long long* keys, int* shifted, int W, int H, hipStream_t s);
void launch_fill_count(const int* seg, uint8_t* border, int* num_neg, int W, int H,
                       hipStream_t s);
void launch_fill_phases(int* seg, const uint8_t* border, const double* centers, int W, int H,
                        hipStream_t s);
void launch_invalidate_disconnected(int* seg, int* parent, int* rsize, unsigned long long* best,
                                    int W, int H, int n, hipStream_t s);
void launch_count_labels(const int* seg, int npix, int* counts, int n, hipStream_t s);
void launch_prior_update(const int* counts_prev, const int* counts_s, const int* counts_f,
                         float* prior_counts, float* invalid_counts, int n, hipStream_t s);
void launch_prop_params_init(Params P, InParams in, int nspix, int nbuf, int sp_size,
                             const int* filled, int npix, hipStream_t s);
void launch_relabel(int* seg, int npix, Params P, int nbuf, const float* prev_prior_mu_app,
                    int nspix_prev, int* max_spix, float thresh_replace, float thresh_new,
                    int* alive, int* target, int* want_new, int* rank, int* src_of, int* total,
                    hipStream_t s);

}  // namespace stsp

namespace stsp {
void launch_marked(const float* vid, const int* spix, const float* color, int T, int H, int W,
                   float* marked, hipStream_t s);
}

namespace stsp {
// k_sweep.hip: returns the buffer that holds the swept labels
int* launch_sweeps(const float* img, int* seg_a, int* seg_b, const SweepRec* rec,
                   int niters_seg, float sigma2_app, float potts, int H, int W,
                   const SimpleLut& lut, int shape, hipStream_t s);
}
