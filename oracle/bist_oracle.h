/*
 * CPU oracle for the streaming BASS/BIST superpixel path -- TEST INFRASTRUCTURE.
 *
 * This is a sequential C restatement of the reference algorithm
 * (st_spix/csrc/bist/ *.cu of gauenk/st_spix), used only by tests/, by the
 * smoke check and by the CPU-baseline leg of bench.py.  It is never linked
 * into, loaded by or called from the product library.  Where the reference
 * is order-nondeterministic (float atomics, first-arrival atomics, atomicAdd
 * ID allocation) the oracle implements the documented deterministic rule
 * that the GPU implementation also follows (rules R1..R15 in the source).
 */
#ifndef BIST_ORACLE_H
#define BIST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int niters;          /* outer iterations (frame 0 / BASS) */
  int niters_seg;      /* boundary sweeps per iteration (reference: 4) */
  int sm_start;        /* split/merge only for idx > sm_start (reference: 0) */
  int sp_size;
  int sm_period;       /* split at idx%p==0, merge at idx%p==p/2 (reference: 4) */
  int prop_niters;     /* >0: fixed iterations for propagated frames; 0: adaptive 12/8/4 */
  float sigma2_app;
  float potts;
  float alpha_hastings; /* log alpha */
  float split_alpha;
  float merge_alpha;
  float thresh_relabel;
  float thresh_new;
  int target_nspix;    /* >0: steer split/merge alphas to this count (bass.cu:65-105) */
  int shape_marginal;  /* 1: add the shape (inverse-Wishart) term to the split ratio
                          (an extension; the reference leaves it unused, 0) */
} orc_cfg;

/* SoA parameter set, field-for-field the reference SuperpixelParams
 * (structs.h:13-48).  Arrays have nspix entries (x3 / x2 where noted). */
typedef struct {
  float* mu_app;             /* 3n */
  float* prior_mu_app;       /* 3n */
  double* mu_shape;          /* 2n */
  double* sigma_shape;       /* 3n, inverse covariance */
  float* logdet_sigma_shape; /* n */
  double* prior_mu_shape;    /* 2n */
  double* prior_sigma_shape; /* 3n */
  double* sample_sigma_shape;/* 3n */
  int* counts;               /* n */
  int* sm_counts;            /* n */
  float* prior_counts;       /* n */
  int* ids;                  /* n */
  float* invalid_counts;     /* n */
} orc_params;

int orc_nspix_from_spsize(int sp_size, int width, int height);
void orc_init_hex_bruteforce(int* seg, int sp_size, int width, int height);
void orc_init_hex(int* seg, int sp_size, int width, int height);
void orc_simple_lut(unsigned char* out256);
double orc_det_log(double x);

/* per-label statistics: out[k*11 + j] = count, S_r, S_g, S_b, X, Y, XX, XY,
 * YY, Q_r, Q_g (Q_b in out2[k]) -- see orc_label_stats for the layout. */
void orc_label_stats(const float* img, const int* seg, int npix, int width,
                     int nlabels, int64_t* out /* nlabels x 12 */);

/* One update_seg call (niters_seg sweeps) with parameters given per label. */
void orc_update_seg(const float* img, int* seg, int height, int width,
                    const float* mu_app, const double* mu_shape,
                    const double* sigma_shape, const double* logdet,
                    const int* counts, int niters_seg, float sigma2_app,
                    float potts);

/* Single-image BASS (run_bass, bass.cu:221-346).  Writes labels and the
 * compacted parameters (capacity cap entries).  Returns nspix or <0. */
int orc_bass(const float* img, int height, int width, const orc_cfg* cfg,
             int* seg_out, orc_params* out, int cap);

/* One propagated frame (shift_and_fill + run_prop, pyapi.cu:128-170).
 * prev: parameters of frame t-1 (nspix_prev entries; prior_counts and
 * invalid_counts are updated in place as the reference does).
 * niters_state: in/out adaptive iteration count (pyapi.cu:141-149).
 * shifted_out / filled_out may be NULL.  Returns nspix or <0. */
int orc_prop(const float* img, const float* flow, int height, int width,
             const int* spix_prev, orc_params* prev, int nspix_prev,
             const orc_cfg* cfg, int* niters_state, int* seg_out,
             orc_params* out, int cap, int* shifted_out, int* filled_out);

/* Pieces of shift_and_fill exposed for unit checks. */
void orc_shift_labels(const int* spix, const float* flow_ds, const int* sizes,
                      int height, int width, int* shifted);
int orc_fill_missing(int* seg, const double* centers, int height, int width,
                     int break_iter);
void orc_invalidate_disconnected(int* seg, int height, int width, int nspix);

/* The per-step ops of the st_spix.prop pipeline (prop_cuda). */
/* shift_labels, csrc/prop/shift_labels.cu:28-62 and 81-126: pixels move by
 * their label's integer flow flow[nspix][2] = (dx, dy); each destination
 * counts arrivals and keeps the largest label (-1 where none).  Labels
 * outside [0, nspix) do not move. */
void orc_prop_shift_labels(const int* spix, const int* flow, int nspix, int height, int width,
                           int* shifted, int* counts);
/* fill_missing, csrc/prop/fill_missing.cu:49-110, 127-140, 142-273, 276-314:
 * fills the -1 pixels listed in missing[nmissing] from their 4-neighbours by
 * the nearest float centre centers[ncenters][2]; labels >= ncenters are not
 * candidates.  Returns the number of rounds run; *negatives receives the
 * listed -1 entries left. */
int orc_prop_fill_missing(int* seg, const float* centers, int ncenters, const int* missing,
                          int nmissing, int height, int width, int break_iter, int* negatives);

/* stsp_shape_marginal of one part (n pixels, position sums x, y, xx, xy,
 * yy), exposed for the closed-form pin */
double orc_shape_marginal(long long n, const long long* sums, const double* psi, double nu);

/* Single steps of the restatement, exposed so the tests can pin each one
 * against closed forms, library routines and brute force. */
/* calc_joint (update_seg.cu:194-259) of one pixel for one candidate label */
float orc_calc_joint(const float* imgC, int x, int y, const float* mu_app,
                     const double* mu_shape, const double* isig, double logdet, int count,
                     float sigma2_app, float neigh_neq, float potts);
/* sum over channels of the NIG marginal terms (split_merge_prop.cu:2934-3070) */
double orc_nig_marginal(float sigma2_app, double n, const double* bn);
/* calc_simple_update (update_params.cu:220-381) for one label; stat =
 * count, S_r, S_g, S_b (2^-32 fixed point), X, Y, XX, XY, YY */
void orc_calc_simple_update(const int64_t* stat, int prop, float prior_count,
                            const double* prior_icov, const double* sample_sigma_shape,
                            int sp_size, float* mu_app, double* mu_shape, double* isig,
                            double* logdet);
/* invert_cov (sparams_io.cu:504-520) */
void orc_invert_cov(const double* cov, double* out);
/* run_sp_downsample (sp_pooling.cu:27-71) as orc_prop computes it */
void orc_flow_pool(const float* flow, const int* seg, int npix, int nspix, float* flow_ds,
                   int* counts);
/* split proposal and Hastings ratios for labels 0..K-1 of seg (all valid,
 * max_spix = K-1): halves map, hasting, accept and select per label */
int orc_split_eval(const float* img, const int* seg, const int* shifted, int height, int width,
                   int K, const int* prop, int direction, float sigma2_app, float split_alpha,
                   int prop_mode, int* halves, double* hasting, int* accept, int* select);
/* merge candidates, Hastings ratios and the winner of each target f */
int orc_merge_eval(const float* img, const int* seg, int height, int width, int K,
                   const int* prop, int direction, float sigma2_app, float alpha_hastings,
                   float merge_alpha, int prop_mode, int* cand, double* hasting, int* winner);
/* relabel targets of labels 0..max_spix; returns the new max label */
int orc_relabel_search(const float* mu_app, const unsigned char* living, int max_spix,
                       const float* prior_mu_app, int nspix_prev, float thresh_relabel,
                       float thresh_new, int nbuf, int* target);

#ifdef __cplusplus
}
#endif
#endif
