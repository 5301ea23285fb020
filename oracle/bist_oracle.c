/*
 * CPU oracle for the streaming BASS/BIST superpixel path.
 *
 * TEST INFRASTRUCTURE ONLY.  A sequential restatement of the reference
 * (gauenk/st_spix, st_spix/csrc/bist/ *.cu); every function cites the
 * reference file:line it follows.  Only tests/, the smoke check and the CPU
 * baseline of bench.py load it, and only as the checker.
 *
 * Deterministic rules where the reference is order-nondeterministic (the GPU
 * implementation follows the same rules):
 *  R1  colour sums, colour square sums and flow sums are exact int64
 *      fixed-point accumulations (colour, squares: 2^-32; flow: 2^-24),
 *      replacing float/double atomicAdd (update_params.cu:114-117,
 *      split_merge_prop.cu:1960-1966, sp_pooling.cu:44-45).
 *  R2  merge: among the k that accept a merge into f, the largest k wins
 *      (the value atomicMax leaves at sm_pairs[2f]; split_merge_prop.cu:
 *      2762-2766 lets the first arrival win).
 *  R3  remove_sp / remove_sp_p read valid/merge/pairs/prior_count from the
 *      state before the step (snapshot semantics).
 *  R4  split: accepted k receive new ids max_spix+1+rank in ascending k
 *      (reference: atomicAdd order, split_merge_prop.cu:3111).
 *  R5  shift collisions: lexicographic min of (size, id) (reference: equal
 *      sizes won by arrival order, shift_labels.cu:111).
 *  R6  invalidate_disconnected: largest 8-connected region per label, ties
 *      to the region whose first pixel (raster order) comes first
 *      (reference: first arrival, atomic_helpers.h:67-83).
 *  R7  relabel: nearest dead candidate ties to the smaller id; new ids in
 *      ascending spix order; when several labels map to one id, parameters
 *      are copied from the largest source label (relabel.cu:85-276).
 *  R8  relabel self-distance of labels born in this frame (>= nspix_prev)
 *      is 0 (reference reads uninitialised memory, relabel.cu:243-245).
 *  R9  split: ninvalid is cleared before each split (never cleared in the
 *      reference, split_merge_prop.cu:461-530).
 *  R10 run_update_prior visits living ids only (reference indexes past the
 *      end of the living-id list, prop.cu:840-846).
 *  R11 icount and sm_count start at 0 (uninitialised in the reference).
 *  R12 merge: the helper b_n of a target f that has no merge candidate of
 *      its own is 0 (init_sm clears it, split_merge.cu:375-377), so its
 *      marginal is +inf and the merge is rejected; the propagated path uses
 *      the same rule (init_sm_p does not clear it, split_merge_prop.cu:461).
 *  R13 NIG log marginals are evaluated in FP64; the lgamma(a_n) terms of
 *      numerator and denominator cancel exactly and are omitted
 *      (split_merge_prop.cu:2988-3019 computes them in FP32).
 *  R14 labels >= 100000 pixels are shifted like any other (reference
 *      initial min-size 100000 drops them, shift_labels.cu:213-216).
 *  R15 logs go through the deterministic stsp_det_log (FP64); the float
 *      narrowing of logdet_sigma_shape follows update_seg.cu:223.
 *  R16 a split whose new id max_spix+1+rank would not fit in the nbuf slots
 *      is dropped (the reference writes past its buffer,
 *      split_merge_prop.cu:3111-3144); relabel never spawns ids >= nbuf.
 */
#include "bist_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../st_spix_amd/csrc/detmath.h"

#define FX_COLOR 4294967296.0 /* 2^32 */
#define FX_FLOW 16777216.0    /* 2^24 */

double orc_det_log(double x) { return stsp_det_log(x); }
double orc_shape_marginal(long long n, const long long* sums, const double* psi, double nu) {
  int64_t s5[5];
  for (int c = 0; c < 5; c++) s5[c] = sums[c];
  return stsp_shape_marginal(n, s5, psi, nu);
}

/* ------------------------------------------------------------------ */
/* simple-point table, seg_utils.h:15-27 (a lookup table that must be  */
/* identical)                                            */
/* ------------------------------------------------------------------ */
static const unsigned char LUT[256] = {
    0, 0, 1, 1, 0, 0, 1, 1, 1, 1, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0,
    0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0,
    1, 1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0,
    0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 1,
    0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 0};

void orc_simple_lut(unsigned char* out256) { memcpy(out256, LUT, 256); }

/* ------------------------------------------------------------------ */
/* internal per-label state (spix_params, structs.h:71-106)             */
/* ------------------------------------------------------------------ */
typedef struct {
  float mu_app[3], prior_mu_app[3];
  double mu_shape[2], sigma_shape[3], prior_mu_shape[2], prior_sigma_shape[3];
  double sample_sigma_shape[3], prior_icov[3];
  double logdet_sigma_shape;
  int count;
  float icount;
  int sm_count;
  float prior_count;
  int valid;
  int prop;
} osp;

/* per-label sums (spix_helper, structs.h:108-113) */
typedef struct {
  int count;
  int64_t S[3], Q[3];
  int64_t X, Y, XX, XY, YY;
  int ninvalid;
} ostat;

static inline int64_t fx_color(float v) {
  return (int64_t)llrint((double)v * FX_COLOR);
}
static inline int64_t fx_square(float v) {
  double d = (double)v;
  return (int64_t)llrint(d * d * FX_COLOR);
}
static inline double fx_to_double(int64_t s) { return (double)s * (1.0 / FX_COLOR); }

static void stats_accumulate(const float* img, const int* seg, int npix, int width,
                             ostat* st, int nst, const int* shifted) {
  for (int t = 0; t < npix; t++) {
    int k = seg[t];
    if (k < 0 || k >= nst) continue;
    const float* p = img + 3 * t;
    ostat* s = st + k;
    s->count += 1;
    for (int c = 0; c < 3; c++) {
      s->S[c] += fx_color(p[c]);
      s->Q[c] += fx_square(p[c]);
    }
    int64_t x = t % width, y = t / width;
    s->X += x;
    s->Y += y;
    s->XX += x * x;
    s->XY += x * y;
    s->YY += y * y;
    if (shifted && shifted[t] < 0) s->ninvalid += 1;
  }
}

void orc_label_stats(const float* img, const int* seg, int npix, int width,
                     int nlabels, int64_t* out) {
  ostat* st = (ostat*)calloc((size_t)nlabels, sizeof(ostat));
  stats_accumulate(img, seg, npix, width, st, nlabels, NULL);
  for (int k = 0; k < nlabels; k++) {
    int64_t* o = out + (size_t)k * 12;
    o[0] = st[k].count;
    o[1] = st[k].S[0]; o[2] = st[k].S[1]; o[3] = st[k].S[2];
    o[4] = st[k].X; o[5] = st[k].Y;
    o[6] = st[k].XX; o[7] = st[k].XY; o[8] = st[k].YY;
    o[9] = st[k].Q[0]; o[10] = st[k].Q[1]; o[11] = st[k].Q[2];
  }
  free(st);
}

/* ------------------------------------------------------------------ */
/* hexagonal initialisation, init_seg.cu:11-81                          */
/* ------------------------------------------------------------------ */
static void hex_geometry(int sp_size, int width, int height, double* Hh, double* w,
                         int* nx, int* ny) {
  double H = sqrt((double)pow((double)sp_size, 2) / (1.5 * sqrt(3.0)));
  double ww = sqrt(3.0) * H;
  *Hh = H;
  *w = ww;
  *nx = (int)floor((double)width / ww) + 2;
  *ny = (int)floor((double)height / (1.5 * H)) + 2;
}

int orc_nspix_from_spsize(int sp_size, int width, int height) {
  double H, w;
  int nx, ny;
  hex_geometry(sp_size, width, height, &H, &w, &nx, &ny);
  return nx * ny;
}

static inline void hex_center(int idx, int nx, double H, double w, double* cx, double* cy) {
  int x = idx % nx, y = idx / nx;
  double xx = (double)x * w;
  double yy = (double)y * 1.5 * H;
  if (y % 2 == 0) xx = xx + 0.5 * w;
  *cx = xx;
  *cy = yy;
}

/* InitHexSeg, init_seg.cu:63-81: nearest centre, ties to the last index. */
void orc_init_hex_bruteforce(int* seg, int sp_size, int width, int height) {
  double H, w;
  int nx, ny;
  hex_geometry(sp_size, width, height, &H, &w, &nx, &ny);
  int K = nx * ny;
  double* c = (double*)malloc(sizeof(double) * 2 * K);
  for (int j = 0; j < K; j++) hex_center(j, nx, H, w, c + 2 * j, c + 2 * j + 1);
  for (int idx = 0; idx < width * height; idx++) {
    int x = idx % width, y = idx / width;
    double D2 = DBL_MAX;
    int best = -1;
    for (int j = 0; j < K; j++) {
      double dx = (x - c[2 * j]), dy = (y - c[2 * j + 1]);
      double d2 = dx * dx + dy * dy;
      if (d2 <= D2) {
        D2 = d2;
        best = j;
      }
    }
    seg[idx] = best;
  }
  free(c);
}

/* Same result as the brute force: the minimum over a 5x5 window of hex
 * cells around the pixel contains every centre at minimal distance. */
void orc_init_hex(int* seg, int sp_size, int width, int height) {
  double H, w;
  int nx, ny;
  hex_geometry(sp_size, width, height, &H, &w, &nx, &ny);
  for (int idx = 0; idx < width * height; idx++) {
    int x = idx % width, y = idx / width;
    int r0 = (int)floor((double)y / (1.5 * H));
    double D2 = DBL_MAX;
    int best = -1;
    for (int r = r0 - 2; r <= r0 + 2; r++) {
      if (r < 0 || r >= ny) continue;
      double off = (r % 2 == 0) ? 0.5 * w : 0.0;
      int c0 = (int)floor(((double)x - off) / w);
      for (int cc = c0 - 2; cc <= c0 + 2; cc++) {
        if (cc < 0 || cc >= nx) continue;
        int j = r * nx + cc;
        double cx, cy;
        hex_center(j, nx, H, w, &cx, &cy);
        double dx = (x - cx), dy = (y - cy);
        double d2 = dx * dx + dy * dy;
        if (d2 < D2 || (d2 == D2 && j > best)) {
          D2 = d2;
          best = j;
        }
      }
    }
    seg[idx] = best;
  }
}

/* ------------------------------------------------------------------ */
/* parameter update, update_params.cu:12-381                            */
/* ------------------------------------------------------------------ */
static void clear_fields(osp* P, int nbuf) { /* update_params.cu:54-88 */
  for (int k = 0; k < nbuf; k++) {
    P[k].count = 0;
    P[k].mu_app[0] = P[k].mu_app[1] = P[k].mu_app[2] = 0.f;
    P[k].mu_shape[0] = P[k].mu_shape[1] = 0.0;
    P[k].sigma_shape[0] = P[k].sigma_shape[1] = P[k].sigma_shape[2] = 0.0;
  }
}

/* _add_sigma_smoothing, update_params.cu:206-217 (count arrives as int) */
static void add_sigma_smoothing(const double* in, int count, float pc, int sp_size,
                                double* out) {
  int nf = 50;
  double total_count = 1. * (float)((float)count + pc * (float)nf);
  out[0] = ((double)(pc * (float)sp_size) + count * in[0]) / (total_count + 3.0);
  out[1] = (count * in[1]) / (total_count + 3.0);
  out[2] = ((double)(pc * (float)sp_size) + count * in[2]) / (total_count + 3.0);
}

/* calc_simple_update, update_params.cu:220-381 */
static void calc_simple_update(osp* p, const ostat* s, int sp_size) {
  if (p->valid == 0) return;
  int count_int = p->count;
  float pc = p->prior_count;
  double count = count_int * 1.0;
  double total_count = (double)count_int + (double)(pc * 50.0f);
  if (count_int <= 0) {
    p->valid = 0;
    return;
  }
  double prior_icov[3], prior_sigma[3];
  if (p->prop) {
    add_sigma_smoothing(p->sample_sigma_shape, (int)pc, pc, sp_size, prior_icov);
  } else {
    prior_icov[0] = p->prior_icov[0];
    prior_icov[1] = p->prior_icov[1];
    prior_icov[2] = p->prior_icov[2];
  }
  double pr_det = prior_icov[0] * prior_icov[2] - prior_icov[1] * prior_icov[1];
  pr_det = sqrt(pr_det);
  prior_sigma[0] = (double)pc / pr_det * prior_icov[0];
  prior_sigma[1] = (double)pc / pr_det * prior_icov[1];
  prior_sigma[2] = (double)pc / pr_det * prior_icov[2];

  for (int c = 0; c < 3; c++) p->mu_app[c] = (float)(fx_to_double(s->S[c]) / count);
  double mx = (double)s->X / count, my = (double)s->Y / count;
  p->mu_shape[0] = mx;
  p->mu_shape[1] = my;
  double sx = (double)s->XX - count * mx * mx;
  double sy = (double)s->XY - count * mx * my;
  double sz = (double)s->YY - count * my * my;
  sx = ((double)pc * prior_sigma[0] + sx) / (total_count + 3.0);
  sy = ((double)pc * prior_sigma[1] + sy) / (total_count + 3.0);
  sz = ((double)pc * prior_sigma[2] + sz) / (total_count + 3.0);
  double det = sx * sz - sy * sy;
  if (det <= 0) {
    sx = sx + 0.00001;
    sz = sz + 0.00001;
    det = sx * sz - sy * sy;
    if (det <= 0) det = 0.00001;
  }
  p->sigma_shape[0] = sz / det;
  p->sigma_shape[1] = -sy / det;
  p->sigma_shape[2] = sx / det;
  p->logdet_sigma_shape = stsp_det_log(det);
}

static void update_params(const float* img, const int* seg, int npix, int width, osp* P,
                          ostat* st, int nbuf, int sp_size) {
  clear_fields(P, nbuf);
  memset(st, 0, sizeof(ostat) * (size_t)nbuf);
  stats_accumulate(img, seg, npix, width, st, nbuf, NULL);
  for (int k = 0; k < nbuf; k++) P[k].count = st[k].count;
  for (int k = 0; k < nbuf; k++) calc_simple_update(P + k, st + k, sp_size);
}

/* calc_summ_stats, update_params.cu:141-203 (first-frame summary) */
static void calc_summ_stats(osp* p, const ostat* s) {
  if (p->valid == 0) return;
  int count_int = p->count;
  double count = count_int * 1.0;
  if (count_int <= 0) {
    p->valid = 0;
    return;
  }
  for (int c = 0; c < 3; c++) p->mu_app[c] = (float)(fx_to_double(s->S[c]) / count);
  double mx = (double)s->X / count, my = (double)s->Y / count;
  p->mu_shape[0] = mx;
  p->mu_shape[1] = my;
  double sx = (double)s->XX / count - mx * mx;
  double sy = (double)s->XY / count - mx * my;
  double sz = (double)s->YY / count - my * my;
  double det = sx * sz - sy * sy;
  if (det <= 0) {
    sx = sx + 0.00001;
    sz = sz + 0.00001;
    det = sx * sz - sy * sy;
    if (det <= 0) det = 0.00001;
  }
  p->sigma_shape[0] = sx;
  p->sigma_shape[1] = sy;
  p->sigma_shape[2] = sz;
  p->logdet_sigma_shape = stsp_det_log(det);
}

/* store_sample_sigma_shape_k, update_params.cu:396-438 */
static void store_sample_sigma_shape(osp* P, const ostat* st, int nbuf) {
  for (int k = 0; k < nbuf; k++) {
    osp* p = P + k;
    if (p->valid == 0) continue;
    double count = p->count * 1.0;
    double mx = (double)st[k].X / count, my = (double)st[k].Y / count;
    p->mu_shape[0] = mx;
    p->mu_shape[1] = my;
    if (count > 0) {
      p->sample_sigma_shape[0] = (double)st[k].XX / count - mx * mx;
      p->sample_sigma_shape[1] = (double)st[k].XY / count - mx * my;
      p->sample_sigma_shape[2] = (double)st[k].YY / count - my * my;
    } else {
      memcpy(p->sample_sigma_shape, p->prior_sigma_shape, sizeof(double) * 3);
    }
  }
}

/* ------------------------------------------------------------------ */
/* boundary update, update_seg.cu:29-311, seg_utils.cu:164-195          */
/* ------------------------------------------------------------------ */
static void find_border_pixels(const int* seg, unsigned char* border, int height,
                               int width) {
  for (int idx = 0; idx < width * height; idx++) {
    int x = idx % width, y = idx / width;
    border[idx] = 0;
    if (y < 1 || x < 1 || y >= height - 1 || x >= width - 1) {
      border[idx] = 1;
      continue;
    }
    int C = seg[idx];
    if (C != seg[idx - width] || C != seg[idx + width] || C != seg[idx + 1] ||
        C != seg[idx - 1])
      border[idx] = 1;
  }
}

/* set_nbrs_v1 + ischangbale_by_nbrs, seg_utils.h:30-79 */
static inline int nbr_check(int NW, int N, int NE, int W, int E, int SW, int S, int SE,
                            int label, int* is_valid) {
  int nb[8] = {NW, N, NE, W, E, SW, S, SE};
  int num = 0, diff = 0;
  for (int i = 7; i >= 0; i--) {
    num <<= 1;
    int same = (label == nb[i]) || (nb[i] == -1);
    if (same)
      num++;
    else
      diff++;
  }
  *is_valid = LUT[num];
  return diff;
}

typedef struct {
  const float* mu_app;    /* 3K */
  const double* mu_shape; /* 2K */
  const double* isig;     /* 3K inverse covariance */
  const double* logdet;   /* K */
  const int* count;       /* K */
} sweep_params;

/* calc_joint, update_seg.cu:194-259 */
static inline void calc_joint(const float* imgC, int x, int y, const sweep_params* sp,
                              int k, float sigma2_app, float neigh_neq, float beta,
                              float* best, int* best_label) {
  float res = -1000.0f;
  int count = sp->count[k];
  const float x0 = imgC[0] - sp->mu_app[3 * k + 0];
  const float x1 = imgC[1] - sp->mu_app[3 * k + 1];
  const float x2 = imgC[2] - sp->mu_app[3 * k + 2];
  const int d0 = (int)((double)x - sp->mu_shape[2 * k + 0]);
  const int d1 = (int)((double)y - sp->mu_shape[2 * k + 1]);
  const float sxx = (float)sp->isig[3 * k + 0];
  const float sxy = (float)sp->isig[3 * k + 1];
  const float syy = (float)sp->isig[3 * k + 2];
  const float ld = (float)sp->logdet[k];
  res = res - (x0 * x0 + x1 * x1 + x2 * x2) / sigma2_app;
  res = res - (float)(d0 * d0) * sxx - (float)(d1 * d1) * syy - (float)(2 * d0 * d1) * sxy;
  res = res - ld;
  res = res - beta * neigh_neq;
  res = res - (float)(100 * (count <= 10));
  if (res > *best) {
    *best = res;
    *best_label = k;
  }
}

/* update_seg_subset, update_seg.cu:29-190 */
static void update_seg_subset(const float* img, int* seg, const unsigned char* border,
                              const sweep_params* sp, float sigma2_app, float potts,
                              int height, int width, int xmod, int ymod) {
  for (int y = ymod; y < height; y += 2) {
    for (int x = xmod; x < width; x += 2) {
      int idx = y * width + x;
      if (!border[idx]) continue;
      int C = seg[idx];
      int N = -1, S = -1, E = -1, W = -1, NE = -1, NW = -1, SE = -1, SW = -1;
      if (x > 0) W = seg[idx - 1];
      if (y > 0) N = seg[idx - width];
      if (x < width - 1) E = seg[idx + 1];
      if (y < height - 1) S = seg[idx + width];
      if (y > 0 && x < width - 1) NE = seg[idx - width + 1];
      if (y > 0 && x > 0) NW = seg[idx - width - 1];
      if (x < width - 1 && y < height - 1) SE = seg[idx + width + 1];
      if (x > 0 && y < height - 1) SW = seg[idx + width - 1];
      int vN, vW, vS, vE;
      int dN = nbr_check(NW, N, NE, W, E, SW, S, SE, N, &vN);
      int dW = nbr_check(NW, N, NE, W, E, SW, S, SE, W, &vW);
      int dS = nbr_check(NW, N, NE, W, E, SW, S, SE, S, &vS);
      int dE = nbr_check(NW, N, NE, W, E, SW, S, SE, E, &vE);
      if (!vN || !vW || !vS || !vE) continue;
      const float* imgC = img + 3 * idx;
      float best = -9999.0f;
      int lab = C;
      if (N >= 0)
        calc_joint(imgC, x, y, sp, N, sigma2_app, (float)dN, potts, &best, &lab);
      if (S != N && S >= 0)
        calc_joint(imgC, x, y, sp, S, sigma2_app, (float)dS, potts, &best, &lab);
      if (W != S && W != N && W >= 0)
        calc_joint(imgC, x, y, sp, W, sigma2_app, (float)dW, potts, &best, &lab);
      if (E != W && E != S && E != N && E >= 0)
        calc_joint(imgC, x, y, sp, E, sigma2_app, (float)dE, potts, &best, &lab);
      if (sp->count[C] <= 2 && lab == C) {
        if (N >= 0) lab = N;
        else if (E >= 0) lab = E;
        else if (S >= 0) lab = S;
        else if (W >= 0) lab = W;
      }
      seg[idx] = lab;
    }
  }
}

static void update_seg_sp(const float* img, int* seg, unsigned char* border,
                          const sweep_params* sp, int niters_seg, float sigma2_app,
                          float potts, int height, int width) {
  for (int it = 0; it < niters_seg; it++) {
    find_border_pixels(seg, border, height, width);
    for (int xm = 0; xm < 2; xm++)
      for (int ym = 0; ym < 2; ym++)
        update_seg_subset(img, seg, border, sp, sigma2_app, potts, height, width, xm, ym);
  }
  find_border_pixels(seg, border, height, width);
}

void orc_update_seg(const float* img, int* seg, int height, int width, const float* mu_app,
                    const double* mu_shape, const double* sigma_shape,
                    const double* logdet, const int* counts, int niters_seg,
                    float sigma2_app, float potts) {
  sweep_params sp = {mu_app, mu_shape, sigma_shape, logdet, counts};
  unsigned char* border = (unsigned char*)malloc((size_t)width * height);
  update_seg_sp(img, seg, border, &sp, niters_seg, sigma2_app, potts, height, width);
  free(border);
}

/* gather the per-label sweep inputs from the AoS state */
typedef struct {
  float* mu_app;
  double* mu_shape;
  double* isig;
  double* logdet;
  int* count;
} sweep_buf;

static void update_seg_P(const float* img, int* seg, unsigned char* border, const osp* P,
                         int nbuf, sweep_buf* b, int niters_seg, float sigma2_app,
                         float potts, int height, int width) {
  for (int k = 0; k < nbuf; k++) {
    memcpy(b->mu_app + 3 * k, P[k].mu_app, sizeof(float) * 3);
    memcpy(b->mu_shape + 2 * k, P[k].mu_shape, sizeof(double) * 2);
    memcpy(b->isig + 3 * k, P[k].sigma_shape, sizeof(double) * 3);
    b->logdet[k] = P[k].logdet_sigma_shape;
    b->count[k] = P[k].count;
  }
  sweep_params sp = {b->mu_app, b->mu_shape, b->isig, b->logdet, b->count};
  update_seg_sp(img, seg, border, &sp, niters_seg, sigma2_app, potts, height, width);
}

/* ------------------------------------------------------------------ */
/* split / merge (single image), split_merge_orig.cu + split_merge.cu   */
/* and the propagated variants, split_merge_prop.cu                     */
/* ------------------------------------------------------------------ */
typedef struct {
  int count, count_f;
  int64_t S[3], Q[3];
  int ninvalid;
  int64_t X[5]; /* position sums x, y, xx, xy, yy (the shape term) */
  double b_n[3], b_n_f[3];
  int has_bn;
  double hasting;
  int merge, remove, select;
} osm;

typedef struct {
  double a0, b0, logb0, logpi, log2;
} nig_consts;

static nig_consts make_nig(float sigma2_app) {
  nig_consts c;
  float a_0 = 1e4f;
  float i_std = sqrtf(sigma2_app) * 2.0f;
  float b_0 = i_std * a_0;
  c.a0 = (double)a_0;
  c.b0 = (double)b_0;
  c.logb0 = stsp_det_log(c.b0);
  c.logpi = stsp_det_log(3.14159265358979323846);
  c.log2 = stsp_det_log(2.0);
  return c;
}

/* b_n = b_0 + 0.5*(sum x^2 - (sum x)^2/n), split_merge_prop.cu:2888-2908 */
static inline double nig_bn(const nig_consts* c, double Q, double S, double n) {
  return c->b0 + 0.5 * (Q - (S * S) / n);
}

/* total marginal sum_c (num - den_c), split_merge_prop.cu:2975-3019, 3052-3070
 * (lgamma terms cancel, rule R13) */
static inline double nig_marginal(const nig_consts* c, double n, const double* bn) {
  double a_n = c->a0 + n / 2.0;
  double num = a_n * c->logb0 + 0.5 * stsp_det_log(1.0 / n);
  double t = 0.5 * n * c->logpi + n * c->log2;
  double m = 0.0;
  for (int ch = 0; ch < 3; ch++) {
    double den = a_n * stsp_det_log(bn[ch]) + t;
    m = m + (num - den);
  }
  return m;
}

/* 4-connected BFS distances from a seed inside its label,
 * calc_split_candidate, split_merge.cu:757-800 (level-synchronous there;
 * a FIFO BFS gives the same distances). */
static void bfs_from_seeds(const int* seg, int* dist, int height, int width, int* queue) {
  int npix = width * height, qh = 0, qt = 0;
  for (int i = 0; i < npix; i++)
    if (dist[i] == 1) queue[qt++] = i;
  while (qh < qt) {
    int i = queue[qh++];
    int x = i % width, y = i / width, L = seg[i], d = dist[i];
    int nb[4] = {y > 0 ? i - width : -1, x > 0 ? i - 1 : -1,
                 y < height - 1 ? i + width : -1, x < width - 1 ? i + 1 : -1};
    for (int j = 0; j < 4; j++) {
      int n = nb[j];
      if (n >= 0 && dist[n] == -1 && seg[n] == L) {
        dist[n] = d + 1;
        queue[qt++] = n;
      }
    }
  }
}

/* init_split / init_split_p, split_merge.cu:803-845 */
static void init_split_seeds(const osp* P, int nbuf, const int* seg, int* dist, int width,
                             int height, int direction) {
  for (int k = 0; k < nbuf; k++) {
    if (P[k].valid == 0) continue;
    int x, y;
    if (direction == 1 || direction == -1) {
      x = (int)(P[k].mu_shape[0]) + direction;
      y = (int)(P[k].mu_shape[1]);
    } else {
      x = (int)(P[k].mu_shape[0]);
      y = (int)(P[k].mu_shape[1]) + direction;
    }
    int ind = y * width + x;
    if (ind < 0 || ind > width * height - 1) continue;
    if (seg[ind] != k) continue;
    dist[ind] = 1;
  }
}

typedef struct {
  int* d1;
  int* d2;
  int* queue;
  int* halves;
  int* pairs;
  osm* sm;
  int nsm;
  int shape_marg; /* add the shape (inverse-Wishart) term to the split ratio */
} sm_work;

/* the two-seed split proposal shared by both paths: returns the half map */
static void split_halves(const int* seg, const osp* P, int nbuf, int max_spix, int direction,
                         int width, int height, sm_work* w) {
  int npix = width * height;
  for (int i = 0; i < npix; i++) w->d1[i] = w->d2[i] = -1;
  init_split_seeds(P, nbuf, seg, w->d1, width, height, direction);
  init_split_seeds(P, nbuf, seg, w->d2, width, height, -direction);
  bfs_from_seeds(seg, w->d1, height, width, w->queue);
  bfs_from_seeds(seg, w->d2, height, width, w->queue);
  /* calc_seg_split, split_merge.cu:847-860 */
  for (int t = 0; t < npix; t++) {
    int v = seg[t];
    if (w->d1[t] > w->d2[t]) v += (max_spix + 1);
    w->halves[t] = v;
  }
}

static void sm_clear(sm_work* w, int nbuf) {
  memset(w->sm, 0, sizeof(osm) * (size_t)w->nsm);
  for (int k = 0; k < w->nsm; k++) w->sm[k].hasting = -999999.0;
  for (int k = 0; k < nbuf; k++) w->pairs[2 * k] = w->pairs[2 * k + 1] = -1;
}

static void sm_accumulate(const float* img, const int* lab, const int* shifted, int npix,
                          int width, sm_work* w) {
  for (int t = 0; t < npix; t++) {
    int k = lab[t];
    if (k < 0 || k >= w->nsm) continue;
    osm* s = w->sm + k;
    s->count += 1;
    for (int c = 0; c < 3; c++) {
      s->S[c] += fx_color(img[3 * t + c]);
      s->Q[c] += fx_square(img[3 * t + c]);
    }
    int64_t x = t % width, y = t / width;
    s->X[0] += x;
    s->X[1] += y;
    s->X[2] += x * x;
    s->X[3] += x * y;
    s->X[4] += y * y;
    if (shifted && shifted[t] < 0) s->ninvalid += 1;
  }
}

/* CudaCalcSplitCandidate_orig / _p: split_merge_orig.cu:202-420,
 * split_merge_prop.cu:184-395.  prop_mode selects update_split_flag_p. */
/* calc_split_stats_step0/1(_p) + the Hastings ratio of update_split_flag(_p):
 * hasting and select per label, acc[k] = 1 where the split is accepted
 * (split_merge_prop.cu:2861-3070, split_merge.cu:862-962) */
static void split_eval(osp* P, const ostat* st, int nbuf, int max_spix, const nig_consts* nc,
                       float split_alpha, int prop_mode, sm_work* w, int* acc) {
  for (int k = 0; k < nbuf; k++) {
    if (P[k].valid == 0) continue;
    int s = k + (max_spix + 1);
    if (s >= nbuf) continue;
    int count_f = P[k].count, count_k = w->sm[k].count, count_s = w->sm[s].count;
    if (count_f < 1 || count_k < 1 || count_s < 1) continue;
    if (count_f != count_k + count_s) continue;
    double bk[3], bs[3], bf[3];
    for (int c = 0; c < 3; c++) {
      double Qk = fx_to_double(w->sm[k].Q[c]), Qs = fx_to_double(w->sm[s].Q[c]);
      bk[c] = nig_bn(nc, Qk, fx_to_double(w->sm[k].S[c]), (double)count_k);
      bs[c] = nig_bn(nc, Qs, fx_to_double(w->sm[s].S[c]), (double)count_s);
      bf[c] = nig_bn(nc, Qk + Qs, fx_to_double(st[k].S[c]), (double)count_f);
    }
    double mk = nig_marginal(nc, (double)count_k, bk);
    double ms = nig_marginal(nc, (double)count_s, bs);
    double mf = nig_marginal(nc, (double)count_f, bf);
    double h = mk + ms - mf + (double)split_alpha;
    if (w->shape_marg)  /* extension, default off: the reference leaves it unused */
      h += stsp_split_shape_term(count_k, w->sm[k].X, count_s, w->sm[s].X,
                                 P[k].prior_sigma_shape, (double)P[k].prior_count);
    if (prop_mode) {
      /* update_split_flag_p, split_merge_prop.cu:3053-3068 */
      int is_prop = P[k].prop;
      int nk = w->sm[k].ninvalid, ns = w->sm[s].ninvalid;
      int propogate_k = (nk <= ns);
      w->sm[k].select = propogate_k || !is_prop;
      double iperc_k = nk / (1.0 * count_k);
      double iperc_s = nk / (1.0 * count_s); /* uses ninvalid_k, as the reference */
      double iperc = propogate_k ? iperc_k : iperc_s;
      iperc = is_prop ? iperc : 1.0;
      h = h - 2.0 * (1.0 - iperc);
    }
    w->sm[k].hasting = h;
    if (h > -2.0) acc[k] = 1;
  }
}

static int split_step(const float* img, int* seg, const int* shifted, osp* P,
                      const ostat* st, int nbuf, int max_spix, int direction,
                      const nig_consts* nc, float split_alpha, int sp_size, int prop_mode,
                      int width, int height, sm_work* w) {
  if (max_spix > nbuf / 2) return max_spix;
  int npix = width * height;
  sm_clear(w, nbuf);
  split_halves(seg, P, nbuf, max_spix, direction, width, height, w);
  sm_accumulate(img, w->halves, prop_mode ? shifted : NULL, npix, width, w);
  int* acc = (int*)calloc((size_t)nbuf, sizeof(int));
  split_eval(P, st, nbuf, max_spix, nc, split_alpha, prop_mode, w, acc);
  (void)sp_size;
  /* rule R4: new ids in ascending k */
  int next = max_spix;
  for (int k = 0; k < nbuf; k++) {
    if (!acc[k]) continue;
    if (next + 1 >= nbuf) { /* no slot left for the new label: dropped */
      acc[k] = 0;
      continue;
    }
    int s = ++next;
    w->pairs[2 * k] = s;
    if (!prop_mode) {
      /* update_split_flag, split_merge_orig.cu:970-977 */
      float pc = (float)fmax(P[k].prior_count / 2.0, 8.0);
      P[k].prior_count = pc;
      P[s].prior_count = pc;
    } else {
      /* update_split_flag_p, split_merge_prop.cu:3108-3144 */
      int nk = w->sm[k].ninvalid, ns = w->sm[k + max_spix + 1].ninvalid;
      int prop_k = (nk <= ns);
      float prior_count = P[k].prior_count;
      float half = (float)fmax(prior_count / 2.0, 36.0);
      if (prop_k) {
        P[k].prior_count = prior_count;
        P[s].prior_count = (float)(prior_count / 2.);
        P[s].prop = 0;
      } else {
        P[k].prior_count = half;
        P[s].prior_count = half;
        P[s].prop = 0;
        P[k].prior_icov[0] = P[s].prior_count;
        P[k].prior_icov[1] = 0;
        P[k].prior_icov[2] = P[s].prior_count;
      }
      P[k].icount = (float)(P[k].icount / 2.0);
      P[s].prior_icov[0] = P[s].prior_count;
      P[s].prior_icov[1] = 0;
      P[s].prior_icov[2] = P[s].prior_count;
      P[s].valid = 1;
    }
  }
  /* split_sp / split_sp_p, split_merge.cu:1002-1033, split_merge_prop.cu:2317-2349 */
  for (int t = 0; t < npix; t++) {
    int k = seg[t];
    if (k < 0 || k >= nbuf || !acc[k]) continue;
    int s = w->pairs[2 * k];
    int k2 = k + (max_spix + 1);
    int take = prop_mode ? (w->sm[k].select ? (w->halves[t] == k2) : (w->halves[t] == k))
                         : (w->halves[t] == k2);
    if (take) {
      seg[t] = s;
      P[s].valid = 1;
    }
  }
  free(acc);
  return next;
}

/* CudaCalcMergeCandidate_orig / _p: split_merge_orig.cu:79-199,
 * split_merge_prop.cu:87-181 */
/* candidates, sums, calc_merge_stats_step0/1/2(_p) and update_merge_flag(_p):
 * pairs[2k+1] = k's candidate f, sm[k].hasting, win[f] = the accepted k that
 * takes f (rule R2) or -1 */
static void merge_eval(const float* img, const int* seg, const unsigned char* border,
                       const osp* P, const ostat* st, int nbuf, int direction,
                       const nig_consts* nc, double log_alpha_f, float merge_alpha,
                       int prop_mode, int width, int height, sm_work* w, int* win) {
  int npix = width * height;
  sm_clear(w, nbuf);
  /* calc_merge_candidate(_p), split_merge.cu:719-755, split_merge_prop.cu:1808-1845 */
  for (int idx = 0; idx < npix; idx++) {
    if (!border[idx]) continue;
    int x = idx % width, y = idx / width;
    int C = seg[idx], Wn = -1;
    if (direction == 1) {
      int ok = prop_mode ? (y > 1 && y < height - 2) : (y >= 0 && y < height - 1);
      if (ok) Wn = seg[idx + width];
    } else {
      int ok = prop_mode ? (x > 1 && x < width - 2) : (x > 0 && x < width);
      if (ok) Wn = seg[idx - 1];
    }
    if (Wn >= 0 && C != Wn) {
      if (prop_mode && (P[C].prop == P[Wn].prop)) continue;
      if (w->pairs[2 * C + 1] < Wn) w->pairs[2 * C + 1] = Wn;
    }
  }
  sm_accumulate(img, seg, NULL, npix, width, w);
  /* step0: b_n of k and of k joined with f */
  for (int k = 0; k < nbuf; k++) {
    if (P[k].valid == 0) continue;
    int f = w->pairs[2 * k + 1];
    if (f < 0) continue;
    int count_f = P[f].count, count_k = P[k].count;
    int count_fk = count_f + count_k;
    w->sm[k].count_f = count_fk;
    for (int c = 0; c < 3; c++) {
      double Qk = fx_to_double(w->sm[k].Q[c]), Qf = fx_to_double(w->sm[f].Q[c]);
      double Sk = fx_to_double(st[k].S[c]), Sf = fx_to_double(st[f].S[c]);
      double b = nig_bn(nc, Qk, Sk, (double)count_k);
      double bf = nc->b0 + 0.5 * ((Qk + Qf) - ((Sf + Sk) * (Sf + Sk)) / (double)count_fk);
      if (b < 0) b = 0.1;
      if (bf < 0) bf = 0.1;
      w->sm[k].b_n[c] = b;
      w->sm[k].b_n_f[c] = bf;
    }
    w->sm[k].has_bn = 1;
  }
  /* step2: hastings (step1's marginals inline) */
  double* mk = (double*)malloc(sizeof(double) * (size_t)nbuf);
  for (int k = 0; k < nbuf; k++) {
    if (P[k].valid == 0) continue;
    mk[k] = nig_marginal(nc, (double)P[k].count, w->sm[k].b_n); /* b_n = 0 unless step0 ran */
  }
  for (int k = 0; k < nbuf; k++) win[k] = -1;
  for (int k = 0; k < nbuf; k++) {
    if (P[k].valid == 0) continue;
    int f = w->pairs[2 * k + 1];
    if (f < 0 || P[f].valid == 0) continue;
    int count_k = P[k].count, count_fk = w->sm[k].count_f;
    if (count_k < 1 || count_fk < 1) continue;
    double mfk = nig_marginal(nc, (double)count_fk, w->sm[k].b_n_f);
    double h = mfk - (log_alpha_f + mk[k] + mk[f]) + (double)merge_alpha;
    w->sm[k].hasting = h;
    /* update_merge_flag(_p), rule R2 */
    if (h > -2.0 && k > win[f]) win[f] = k;
  }
  free(mk);
}

static void merge_step(const float* img, int* seg, const unsigned char* border, osp* P,
                       const ostat* st, int nbuf, int direction, const nig_consts* nc,
                       double log_alpha_f, float merge_alpha, int prop_mode, int width,
                       int height, sm_work* w) {
  int npix = width * height;
  int* win = (int*)malloc(sizeof(int) * (size_t)nbuf);
  merge_eval(img, seg, border, P, st, nbuf, direction, nc, log_alpha_f, merge_alpha, prop_mode,
             width, height, w, win);
  for (int f = 0; f < nbuf; f++)
    if (win[f] >= 0) {
      w->sm[win[f]].merge = 1;
      w->pairs[2 * f] = win[f];
    }
  /* remove_sp(_p), split_merge.cu:1035-1080, split_merge_prop.cu:2353-2449 (rule R3) */
  int* pairs0 = (int*)malloc(sizeof(int) * 2 * (size_t)nbuf);
  memcpy(pairs0, w->pairs, sizeof(int) * 2 * (size_t)nbuf);
  int* valid0 = (int*)malloc(sizeof(int) * (size_t)nbuf);
  float* pc0 = (float*)malloc(sizeof(float) * (size_t)nbuf);
  for (int k = 0; k < nbuf; k++) {
    valid0[k] = P[k].valid;
    pc0[k] = P[k].prior_count;
  }
  for (int k = 0; k < nbuf; k++) {
    int s = pairs0[2 * k + 1];
    if (s < 0) continue;
    int is_cycle = pairs0[2 * s + 1] == k;
    if (valid0[k] == 0 || valid0[s] == 0) continue;
    int mk_ = w->sm[k].merge, ms_ = w->sm[s].merge;
    if (!prop_mode) {
      if (mk_ && (!ms_ || (is_cycle && k < s))) {
        w->sm[k].remove = 1;
        P[k].valid = 0;
        P[s].prior_count = pc0[k] + pc0[s];
      }
    } else {
      int is_prop_k = P[k].prop;
      int cond_c = is_cycle && is_prop_k;
      if (mk_ && (!ms_ || cond_c)) {
        int victim = k, keeper = s;
        if (is_prop_k) {
          w->pairs[2 * s + 1] = k;
          victim = s;
          keeper = k;
        }
        w->sm[victim].remove = 1;
        P[victim].valid = 0;
        P[keeper].prior_count = pc0[victim] + pc0[keeper];
      }
    }
  }
  /* merge_sp(_p), split_merge.cu:984-1000 */
  for (int idx = 0; idx < npix; idx++) {
    int k = seg[idx];
    if (k >= 0 && k < nbuf && w->sm[k].remove) seg[idx] = w->pairs[2 * k + 1];
  }
  free(pairs0);
  free(valid0);
  free(pc0);
  free(win);
}

/* ------------------------------------------------------------------ */
/* compaction, compact_spix.cu:41-122, 196-295                          */
/* ------------------------------------------------------------------ */
static void copy_compact_fields(osp* dst, const osp* src) {
  /* fill_new_params_from_old: the fields copied there; icount, sm_count,
   * prop and prior_icov stay with the destination slot. */
  memcpy(dst->mu_app, src->mu_app, sizeof(dst->mu_app));
  memcpy(dst->prior_mu_app, src->prior_mu_app, sizeof(dst->prior_mu_app));
  memcpy(dst->mu_shape, src->mu_shape, sizeof(dst->mu_shape));
  memcpy(dst->sigma_shape, src->sigma_shape, sizeof(dst->sigma_shape));
  memcpy(dst->prior_mu_shape, src->prior_mu_shape, sizeof(dst->prior_mu_shape));
  memcpy(dst->prior_sigma_shape, src->prior_sigma_shape, sizeof(dst->prior_sigma_shape));
  memcpy(dst->sample_sigma_shape, src->sample_sigma_shape,
         sizeof(dst->sample_sigma_shape));
  dst->logdet_sigma_shape = src->logdet_sigma_shape;
  dst->count = src->count;
  dst->prior_count = src->prior_count;
  dst->valid = src->valid;
}

/* living ids >= prev_nspix are renumbered prev_nspix + rank (ascending).
 * Returns prev_nspix + number of such ids. */
static int compactify(int* seg, int npix, osp* P, int nbuf, int prev_nspix, int maxlab) {
  int n = maxlab + 1;
  if (n < 1) n = 1;
  unsigned char* alive = (unsigned char*)calloc((size_t)n, 1);
  for (int i = 0; i < npix; i++)
    if (seg[i] >= 0 && seg[i] < n) alive[seg[i]] = 1;
  int* map = (int*)malloc(sizeof(int) * (size_t)n);
  int num_new = 0;
  for (int k = 0; k < n; k++) {
    map[k] = k;
    if (alive[k] && k >= prev_nspix) map[k] = prev_nspix + num_new++;
  }
  for (int i = 0; i < npix; i++)
    if (seg[i] >= prev_nspix) seg[i] = map[seg[i]];
  if (P && num_new > 0) {
    osp* tmp = (osp*)malloc(sizeof(osp) * (size_t)num_new);
    for (int k = prev_nspix; k < n; k++)
      if (alive[k]) {
        int d = map[k] - prev_nspix;
        tmp[d] = P[k];
      }
    for (int d = 0; d < num_new; d++)
      if (prev_nspix + d < nbuf) copy_compact_fields(P + prev_nspix + d, tmp + d);
    free(tmp);
  }
  free(alive);
  free(map);
  return prev_nspix + num_new;
}

/* Number of distinct labels in seg (extract_unique_ids(seg).size(),
 * compact_spix.cu:196-230, as used at bass.cu:75 and prop.cu:373). */
static int count_living(const int* seg, int npix, int nbuf) {
  unsigned char* seen = (unsigned char*)calloc((size_t)nbuf, 1);
  int n = 0;
  for (int t = 0; t < npix; t++) {
    int k = seg[t];
    if (k >= 0 && k < nbuf && !seen[k]) {
      seen[k] = 1;
      n++;
    }
  }
  free(seen);
  return n;
}

/* Split/merge alpha control toward target_nspix, bass.cu:73-98 and
 * prop.cu:371-400 (the prop variant also doubles thresh_new every 200
 * iterations past 2000).  Returns 1 when the count is accepted. */
static int control_nspix(int nliving, int target, int idx, int og_niters, int prop_mode,
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

static int max_label(const int* seg, int npix) {
  int m = -1;
  for (int i = 0; i < npix; i++)
    if (seg[i] > m) m = seg[i];
  return m;
}

/* get_params_as_vectors + read_params, sparams_io.cu:65-79, 189-286 */
static void export_params(const osp* P, int nspix, orc_params* out) {
  for (int ix = 0; ix < nspix; ix++) {
    const osp* p = P + ix;
    for (int c = 0; c < 3; c++) {
      out->mu_app[3 * ix + c] = p->mu_app[c];
      out->prior_mu_app[3 * ix + c] = p->prior_mu_app[c];
      out->sigma_shape[3 * ix + c] = p->sigma_shape[c];
      out->prior_sigma_shape[3 * ix + c] = p->prior_sigma_shape[c];
      out->sample_sigma_shape[3 * ix + c] = p->sample_sigma_shape[c];
    }
    for (int c = 0; c < 2; c++) {
      out->mu_shape[2 * ix + c] = p->mu_shape[c];
      out->prior_mu_shape[2 * ix + c] = p->prior_mu_shape[c];
    }
    out->logdet_sigma_shape[ix] = (float)p->logdet_sigma_shape;
    out->counts[ix] = p->count;
    out->invalid_counts[ix] = p->icount;
    out->sm_counts[ix] = p->sm_count;
    out->prior_counts[ix] = p->prior_count;
    out->ids[ix] = ix;
  }
}

/* update_prior_kernel (invert=false), sparams_io.cu:583-639; rule R10 */
static void update_prior(orc_params* out, const int* seg, int npix, int nspix) {
  unsigned char* alive = (unsigned char*)calloc((size_t)(nspix > 0 ? nspix : 1), 1);
  for (int i = 0; i < npix; i++)
    if (seg[i] >= 0 && seg[i] < nspix) alive[seg[i]] = 1;
  for (int k = 0; k < nspix; k++) {
    if (!alive[k] || out->counts[k] <= 0) continue;
    for (int c = 0; c < 3; c++) out->prior_mu_app[3 * k + c] = out->mu_app[3 * k + c];
    for (int c = 0; c < 2; c++) out->prior_mu_shape[2 * k + c] = out->mu_shape[2 * k + c];
    for (int c = 0; c < 3; c++)
      out->prior_sigma_shape[3 * k + c] = out->sigma_shape[3 * k + c];
  }
  free(alive);
}

static sm_work sm_alloc(int npix, int nbuf) {
  sm_work w;
  w.d1 = (int*)malloc(sizeof(int) * (size_t)npix);
  w.d2 = (int*)malloc(sizeof(int) * (size_t)npix);
  w.queue = (int*)malloc(sizeof(int) * (size_t)npix);
  w.halves = (int*)malloc(sizeof(int) * (size_t)npix);
  w.pairs = (int*)malloc(sizeof(int) * 2 * (size_t)(nbuf + 2));
  w.nsm = 2 * nbuf + 2;
  w.sm = (osm*)calloc((size_t)w.nsm, sizeof(osm));
  w.shape_marg = 0;
  return w;
}
static void sm_free(sm_work* w) {
  free(w->d1);
  free(w->d2);
  free(w->queue);
  free(w->halves);
  free(w->pairs);
  free(w->sm);
}
static sweep_buf sweep_alloc(int nbuf) {
  sweep_buf b;
  b.mu_app = (float*)malloc(sizeof(float) * 3 * (size_t)nbuf);
  b.mu_shape = (double*)malloc(sizeof(double) * 2 * (size_t)nbuf);
  b.isig = (double*)malloc(sizeof(double) * 3 * (size_t)nbuf);
  b.logdet = (double*)malloc(sizeof(double) * (size_t)nbuf);
  b.count = (int*)malloc(sizeof(int) * (size_t)nbuf);
  return b;
}
static void sweep_free(sweep_buf* b) {
  free(b->mu_app);
  free(b->mu_shape);
  free(b->isig);
  free(b->logdet);
  free(b->count);
}

static inline int is_split_iter(int idx, const orc_cfg* cfg) {
  int p = cfg->sm_period > 0 ? cfg->sm_period : 4;
  return idx > cfg->sm_start && (idx % p) == 0;
}
static inline int is_merge_iter(int idx, const orc_cfg* cfg) {
  int p = cfg->sm_period > 0 ? cfg->sm_period : 4;
  return idx > cfg->sm_start && (idx % p) == p / 2;
}

/* ------------------------------------------------------------------ */
/* run_bass, bass.cu:221-346 with the loop of bass(), bass.cu:46-181    */
/* ------------------------------------------------------------------ */
int orc_bass(const float* img, int height, int width, const orc_cfg* cfg, int* seg,
             orc_params* out, int cap) {
  int npix = height * width, sp_size = cfg->sp_size;
  orc_init_hex(seg, sp_size, width, height);
  int K0 = orc_nspix_from_spsize(sp_size, width, height);
  int nbuf = K0 * 30; /* bass.cu:245, before the first compaction */
  int nspix = compactify(seg, npix, NULL, 0, 0, K0 - 1);
  osp* P = (osp*)calloc((size_t)nbuf, sizeof(osp));
  ostat* st = (ostat*)calloc((size_t)nbuf, sizeof(ostat));
  unsigned char* border = (unsigned char*)calloc((size_t)npix, 1);
  sm_work w = sm_alloc(npix, nbuf);
  w.shape_marg = cfg->shape_marginal;
  sweep_buf sb = sweep_alloc(nbuf);
  /* mark_active_contiguous, init_sparams.cu:195-231 */
  for (int k = 0; k < nbuf; k++) {
    P[k].valid = 0;
    P[k].prop = 0;
    if (k >= nspix) {
      P[k].prior_sigma_shape[0] = 1.0 / sp_size;
      P[k].prior_sigma_shape[1] = 0;
      P[k].prior_sigma_shape[2] = 1.0 / sp_size;
    }
  }
  for (int k = 0; k < nspix; k++) {
    P[k].valid = 1;
    P[k].prop = 1;
  }
  /* init_sp_params, init_sparams.cu:25-166 */
  clear_fields(P, nbuf);
  memset(st, 0, sizeof(ostat) * (size_t)nbuf);
  stats_accumulate(img, seg, npix, width, st, nbuf, NULL);
  for (int k = 0; k < nbuf; k++) P[k].count = st[k].count;
  for (int k = 0; k < nbuf; k++) calc_summ_stats(P + k, st + k);
  int count = (int)(npix / (1. * nspix));
  for (int k = 0; k < nbuf; k++) {
    osp* p = P + k;
    p->prior_sigma_shape[0] = 1.0 / sp_size;
    p->prior_sigma_shape[1] = 0.;
    p->prior_sigma_shape[2] = 1.0 / sp_size;
    if (k < nspix) {
      p->valid = 1;
      p->prior_count = (float)count;
      memcpy(p->prior_mu_app, p->mu_app, sizeof(p->mu_app));
      p->mu_app[0] = p->mu_app[1] = p->mu_app[2] = 0;
      p->prior_icov[0] = 1;
      p->prior_icov[1] = 0;
      p->prior_icov[2] = 1;
      p->mu_shape[0] = p->mu_shape[1] = 0;
      p->sigma_shape[0] = p->sigma_shape[1] = p->sigma_shape[2] = 0;
      p->prop = 0;
    } else {
      p->count = 0;
      memset(p->mu_app, 0, sizeof(p->mu_app));
      memset(p->prior_mu_app, 0, sizeof(p->prior_mu_app));
      p->mu_shape[0] = p->mu_shape[1] = 0;
      p->prior_mu_shape[0] = p->prior_mu_shape[1] = 0;
      p->sigma_shape[0] = p->sigma_shape[1] = p->sigma_shape[2] = 0;
      p->logdet_sigma_shape = 0;
      p->prop = 0;
      p->valid = 0;
      p->prior_icov[0] = 1;
      p->prior_icov[1] = 0;
      p->prior_icov[2] = 1;
    }
  }
  find_border_pixels(seg, border, height, width);

  nig_consts nc = make_nig(cfg->sigma2_app);
  double log_alpha = stsp_det_log((double)expf(cfg->alpha_hastings));
  int smc = 1; /* split/merge counter, bass.cu:56 */
  int max_spix = nspix - 1;
  /* bass.cu:60-67: merge_alpha is local (0) in BASS; 5000-iteration cap
   * when the count is controlled */
  float split_alpha = cfg->split_alpha, merge_alpha = 0.0f, thresh_new = cfg->thresh_new;
  const int controlled = cfg->target_nspix > 0;
  const int niters = controlled ? 5000 : cfg->niters;
  for (int idx = 0; idx < niters; idx++) {
    if (controlled && idx % 2 == 0 &&
        control_nspix(count_living(seg, npix, nbuf), cfg->target_nspix, idx, cfg->niters, 0,
                      &split_alpha, &merge_alpha, &thresh_new))
      break;
    update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    if (is_split_iter(idx, cfg)) {
      smc += 1; /* run_split_orig, split_merge_orig.cu:37-52 */
      int direction = smc % 2 + 1;
      max_spix = split_step(img, seg, NULL, P, st, nbuf, max_spix, direction, &nc,
                            split_alpha, sp_size, 0, width, height, &w);
      update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    }
    if (is_merge_iter(idx, cfg)) {
      int direction = smc % 2; /* run_merge_orig, split_merge_orig.cu:67-75 */
      merge_step(img, seg, border, P, st, nbuf, direction, &nc, log_alpha, merge_alpha, 0, width,
                 height, &w);
      update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    }
    update_seg_P(img, seg, border, P, nbuf, &sb, cfg->niters_seg, cfg->sigma2_app,
                 cfg->potts, height, width);
  }
  update_params(img, seg, npix, width, P, st, nbuf, sp_size);
  store_sample_sigma_shape(P, st, nbuf);
  nspix = compactify(seg, npix, P, nbuf, 0, max_spix > max_label(seg, npix) ? max_spix
                                                                            : max_label(seg, npix));
  int rc = nspix;
  if (nspix > cap) {
    rc = -2;
  } else {
    export_params(P, nspix, out);
    update_prior(out, seg, npix, nspix);
  }
  sweep_free(&sb);
  sm_free(&w);
  free(border);
  free(st);
  free(P);
  return rc;
}

/* propagated frames are restated in bist_oracle_prop.c */

/* ================================================================== */
/* propagated frames: shift_and_fill.cu + prop.cu                       */
/* ================================================================== */

/* shift_labels_kernel + decode_kernel, shift_labels.cu:32-164 (rule R5) */
void orc_shift_labels(const int* spix, const float* flow_ds, const int* sizes, int height,
                      int width, int* shifted) {
  int npix = height * width;
  int* best_size = (int*)malloc(sizeof(int) * (size_t)npix);
  for (int i = 0; i < npix; i++) {
    shifted[i] = -1;
    best_size[i] = 0;
  }
  for (int t = 0; t < npix; t++) {
    int L = spix[t];
    if (L < 0) continue;
    int h = t / width, w = t % width;
    int fw = (int)roundf(flow_ds[2 * L]);
    int fh = (int)roundf(flow_ds[2 * L + 1]);
    int wd = w + fw, hd = h + fh;
    if (wd < 0 || wd >= width || hd < 0 || hd >= height) continue;
    int d = hd * width + wd;
    int sz = sizes[L];
    if (shifted[d] < 0 || sz < best_size[d] || (sz == best_size[d] && L < shifted[d])) {
      shifted[d] = L;
      best_size[d] = sz;
    }
  }
  free(best_size);
}

/* fill_missing + update_missing_seg_nn + find_border_along_missing,
 * fill_missing.cu:35-130, 288-441, 617-647.  Returns 0, or -3 when the
 * number of missing pixels stops shrinking (the reference exits). */
int orc_fill_missing(int* seg, const double* centers, int height, int width,
                     int break_iter) {
  int npix = height * width;
  unsigned char* border = (unsigned char*)malloc((size_t)npix);
  long long prev_neg = -1;
  int iter = 0, rc = 0;
  while (1) {
    if (break_iter > 0 && iter >= break_iter) break;
    int num_neg = 0;
    for (int idx = 0; idx < npix; idx++) {
      int x = idx % width, y = idx / width;
      int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
      int S = y < height - 1 ? seg[idx + width] : -1, E = x < width - 1 ? seg[idx + 1] : -1;
      border[idx] = (N >= 0) || (Wl >= 0) || (S >= 0) || (E >= 0);
      if (seg[idx] < 0) num_neg++;
    }
    if (num_neg == 0) break;
    if (iter > 0 && num_neg == prev_neg) {
      rc = -3;
      break;
    }
    for (int xm = 0; xm < 2; xm++)
      for (int ym = 0; ym < 2; ym++)
        for (int y = ym; y < height; y += 2)
          for (int x = xm; x < width; x += 2) {
            int idx = y * width + x;
            if (!border[idx] || seg[idx] != -1) continue;
            int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
            int S = y < height - 1 ? seg[idx + width] : -1;
            int E = x < width - 1 ? seg[idx + 1] : -1;
            float best = -999999.0f;
            int lab = -1;
            int cand[4] = {N, S, Wl, E};
            for (int j = 0; j < 4; j++) {
              int L = cand[j];
              if (L < 0) continue;
              int dup = 0;
              for (int q = 0; q < j; q++)
                if (cand[q] == L) dup = 1;
              if (dup) continue;
              /* isotropic_space, fill_missing.cu:140-153 */
              float dx = (float)(((double)(1.0f * (float)x) - centers[2 * L]) /
                                 (double)(1.0f * (float)width));
              float dy = (float)(((double)(1.0f * (float)y) - centers[2 * L + 1]) /
                                 (double)(1.0f * (float)height));
              float sim = -dx * dx - dy * dy;
              if (sim > best) {
                best = sim;
                lab = L;
              }
            }
            seg[idx] = lab;
          }
    iter++;
    prev_neg = num_neg;
  }
  free(border);
  return rc;
}

/* prop_cuda.shift_labels: shift_labels_kernel (shift_labels.cu:28-62) for
 * one image; the counts and the atomicMax of the label are order-free. */
void orc_prop_shift_labels(const int* spix, const int* flow, int nspix, int height, int width,
                           int* shifted, int* counts) {
  int npix = height * width;
  for (int i = 0; i < npix; i++) {
    shifted[i] = -1; /* run_shift_labels: -torch::ones */
    counts[i] = 0;
  }
  for (int i = 0; i < npix; i++) {
    int L = spix[i];
    if (L < 0 || L >= nspix) continue;
    int hd = i / width + flow[2 * L + 1], wd = i % width + flow[2 * L];
    if (hd < 0 || hd >= height || wd < 0 || wd >= width) continue;
    int d = hd * width + wd;
    counts[d] += 1;
    if (L > shifted[d]) shifted[d] = L;
  }
}

/* prop_cuda.fill_missing: the host loop of fill_missing (fill_missing.cu:
 * 49-110), find_border_along_missing (276-314) over the listed pixels and
 * four checkerboard passes of update_missing_seg_nn (142-273) with the float
 * isotropic_space (127-140). */
int orc_prop_fill_missing(int* seg, const float* centers, int ncenters, const int* missing,
                          int nmissing, int height, int width, int break_iter, int* negatives) {
  int npix = height * width;
  unsigned char* border = (unsigned char*)calloc((size_t)npix, 1);
  int prev_neg = 1, num_neg = 1, iter = 0;
  while (num_neg > 0) {
    if (break_iter > 0 && iter >= break_iter) break;
    memset(border, 0, (size_t)npix);
    num_neg = 0;
    for (int j = 0; j < nmissing; j++) {
      int idx = missing[j];
      if (idx < 0 || idx >= npix) continue; /* outside the image: ignored */
      int x = idx % width, y = idx / width;
      int C = seg[idx];
      int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
      int S = y < height - 1 ? seg[idx + width] : -1, E = x < width - 1 ? seg[idx + 1] : -1;
      int valid = (N >= 0) || (Wl >= 0) || (S >= 0) || (E >= 0);
      if (valid && C < 0) border[idx] = 1;
      if (C < 0) num_neg++;
    }
    for (int xm = 0; xm < 2; xm++)
      for (int ym = 0; ym < 2; ym++)
        for (int idx = 0; idx < npix; idx++) {
          int x = idx % width, y = idx / width;
          if (x % 2 != xm || y % 2 != ym || !border[idx]) continue;
          int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
          int E = x < width - 1 ? seg[idx + 1] : -1, S = y < height - 1 ? seg[idx + width] : -1;
          float best = -999999.0f, lab = -1.0f; /* res_max (x, y) of float2 */
          int cand[4] = {N, S, Wl, E};
          for (int j = 0; j < 4; j++) {
            int L = cand[j], dup = 0;
            for (int q = 0; q < j; q++) dup |= cand[q] == L;
            if (L < 0 || dup || L >= ncenters) continue;
            float dx = (1.0f * (float)x - centers[2 * L]) / (1.0f * (float)width);
            float dy = (1.0f * (float)y - centers[2 * L + 1]) / (1.0f * (float)height);
            float sim = -dx * dx - dy * dy;
            if (sim > best) {
              best = sim;
              lab = (float)L;
            }
          }
          seg[idx] = (int)lab;
        }
    iter++;
    if (iter > 0 && num_neg == prev_neg) break; /* "the border won't shrink" */
    prev_neg = num_neg;
  }
  int left = 0;
  for (int j = 0; j < nmissing; j++)
    left += missing[j] >= 0 && missing[j] < npix && seg[missing[j]] < 0;
  if (negatives) *negatives = left;
  free(border);
  return iter;
}

/* run_invalidate_disconnected, split_disconnected.cu:33-185 (rule R6) */
void orc_invalidate_disconnected(int* seg, int height, int width, int nspix) {
  int npix = height * width;
  int* comp = (int*)malloc(sizeof(int) * (size_t)npix);
  int* queue = (int*)malloc(sizeof(int) * (size_t)npix);
  for (int i = 0; i < npix; i++) comp[i] = -1;
  int nreg = 0;
  int* reg_size = (int*)malloc(sizeof(int) * (size_t)npix);
  int* reg_lab = (int*)malloc(sizeof(int) * (size_t)npix);
  for (int i = 0; i < npix; i++) {
    if (comp[i] >= 0) continue;
    int L = seg[i], qh = 0, qt = 0;
    comp[i] = nreg;
    queue[qt++] = i;
    while (qh < qt) {
      int p = queue[qh++];
      int px = p % width, py = p / width;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int nx = px + dx, ny = py + dy;
          if (nx < 0 || ny < 0 || nx >= width || ny >= height) continue;
          int q = ny * width + nx;
          if (comp[q] < 0 && seg[q] == L) {
            comp[q] = nreg;
            queue[qt++] = q;
          }
        }
    }
    reg_size[nreg] = qt;
    reg_lab[nreg] = L;
    nreg++;
  }
  /* regions are numbered by their first pixel in raster order */
  int* keep = (int*)malloc(sizeof(int) * (size_t)(nspix > 0 ? nspix : 1));
  int* keep_size = (int*)malloc(sizeof(int) * (size_t)(nspix > 0 ? nspix : 1));
  for (int k = 0; k < nspix; k++) {
    keep[k] = -1;
    keep_size[k] = 0;
  }
  for (int r = 0; r < nreg; r++) {
    int L = reg_lab[r];
    if (L < 0 || L >= nspix) continue;
    if (reg_size[r] > keep_size[L]) {
      keep_size[L] = reg_size[r];
      keep[L] = r;
    }
  }
  for (int i = 0; i < npix; i++) {
    int L = seg[i];
    if (L < 0 || L >= nspix) continue;
    if (keep[L] != comp[i]) seg[i] = -1;
  }
  free(comp);
  free(queue);
  free(reg_size);
  free(reg_lab);
  free(keep);
  free(keep_size);
}

/* relabel_spix, relabel.cu:279-409 (rules R7, R8) */
/* find_most_similar_spix + mark_for_relabel (relabel.cu:125-276, rules R7,
 * R8): target[s] for every label s < max_spix + 1; returns the new max label */
static int relabel_search(const float* mu_app, int mu_stride, const unsigned char* living,
                          int max_spix, const float* prior_mu_app, int nspix_prev,
                          float thresh_replace, float thresh_new, int nbuf, int* target) {
  int nspix = max_spix + 1;
  for (int s = 0; s < nspix; s++) target[s] = -1;
  int next = max_spix;
  for (int s = 0; s < nspix; s++) {
    if (!living[s]) continue;
    /* find_most_similar_spix, relabel.cu:200-276 */
    float best = 10000.0f;
    int cand = -1;
    float ss = 0.0f;
    for (int c = 0; c < nspix_prev; c++) {
      int is_same = (c == s);
      if (living[c] && !is_same) continue;
      const float* m = mu_app + (size_t)mu_stride * s;
      float dx = m[0] - prior_mu_app[3 * c + 0];
      float dy = m[1] - prior_mu_app[3 * c + 1];
      float dz = m[2] - prior_mu_app[3 * c + 2];
      float delta = (dx * dx + dy * dy + dz * dz);
      if (is_same) ss = delta;
      if (delta < best) {
        best = delta;
        cand = c;
      }
    }
    /* mark_for_relabel, relabel.cu:125-192 */
    int revive = (cand >= 0) && (best < thresh_replace) && (cand < nspix) && (cand != s);
    if (revive) {
      target[s] = cand;
    } else if (ss > thresh_new) {
      if (next + 1 < nbuf) target[s] = ++next;
    }
  }
  return next;
}

static int relabel_step(int* seg, int npix, osp* P, int nbuf, const orc_params* prev,
                        float thresh_replace, float thresh_new, int nspix_prev,
                        int max_spix) {
  int nspix = max_spix + 1;
  unsigned char* living = (unsigned char*)calloc((size_t)nspix, 1);
  for (int i = 0; i < npix; i++)
    if (seg[i] >= 0 && seg[i] < nspix) living[seg[i]] = 1;
  int* target = (int*)malloc(sizeof(int) * (size_t)nspix);
  float* mu = (float*)malloc(sizeof(float) * 3 * (size_t)nspix);
  for (int k = 0; k < nspix; k++) memcpy(mu + 3 * k, P[k].mu_app, sizeof(float) * 3);
  int next = relabel_search(mu, 3, living, max_spix, prev->prior_mu_app, nspix_prev,
                            thresh_replace, thresh_new, nbuf, target);
  free(mu);
  /* relabel_spix_kernel, relabel.cu:85-120: parameters from the largest source */
  for (int s = 0; s < nspix; s++) {
    int n = target[s];
    if (n < 0) continue;
    P[n].count = P[s].count;
    P[n].icount = P[s].icount;
    if (n >= nspix_prev) {
      P[n].prop = 0;
      P[n].prior_count = P[s].prior_count;
      memcpy(P[n].prior_sigma_shape, P[s].prior_sigma_shape, sizeof(double) * 3);
      memcpy(P[n].sample_sigma_shape, P[s].sample_sigma_shape, sizeof(double) * 3);
      memcpy(P[n].prior_mu_app, P[s].prior_mu_app, sizeof(float) * 3);
      memcpy(P[n].prior_icov, P[s].prior_icov, sizeof(double) * 3);
    }
  }
  for (int s = 0; s < nspix; s++)
    if (target[s] >= 0) P[s].valid = 0;
  for (int s = 0; s < nspix; s++)
    if (target[s] >= 0) P[target[s]].valid = 1;
  for (int i = 0; i < npix; i++) {
    int L = seg[i];
    if (L >= 0 && L < nspix && target[L] >= 0) seg[i] = target[L];
  }
  free(living);
  free(target);
  return next;
}

int orc_prop(const float* img, const float* flow, int height, int width,
             const int* spix_prev, orc_params* prev, int nspix_prev, const orc_cfg* cfg,
             int* niters_state, int* seg, orc_params* out, int cap, int* shifted_out,
             int* filled_out) {
  int npix = height * width, sp_size = cfg->sp_size, nspix = nspix_prev;
  /* run_sp_downsample, sp_pooling.cu:27-71, 121-147 (rule R1) */
  int64_t* fsum = (int64_t*)calloc((size_t)2 * nspix, sizeof(int64_t));
  int* counts = (int*)calloc((size_t)nspix, sizeof(int));
  for (int t = 0; t < npix; t++) {
    int L = spix_prev[t];
    if (L < 0 || L >= nspix) continue;
    fsum[2 * L] += (int64_t)llrint((double)flow[2 * t] * FX_FLOW);
    fsum[2 * L + 1] += (int64_t)llrint((double)flow[2 * t + 1] * FX_FLOW);
    counts[L]++;
  }
  float* flow_ds = (float*)calloc((size_t)2 * nspix + 2, sizeof(float));
  double* centers = (double*)malloc(sizeof(double) * 2 * (size_t)nspix);
  for (int k = 0; k < nspix; k++) {
    for (int c = 0; c < 2; c++) {
      double sum = (double)fsum[2 * k + c] * (1.0 / FX_FLOW);
      flow_ds[2 * k + c] = counts[k] > 0 ? (float)(sum / (1. * counts[k])) : 0.0f;
      centers[2 * k + c] = prev->mu_shape[2 * k + c] + (double)flow_ds[2 * k + c];
    }
  }
  int* shifted = (int*)malloc(sizeof(int) * (size_t)npix);
  orc_shift_labels(spix_prev, flow_ds, counts, height, width, shifted);
  int* filled = seg; /* the filled map becomes the initial segmentation */
  memcpy(filled, shifted, sizeof(int) * (size_t)npix);
  int* counts_s = (int*)calloc((size_t)nspix, sizeof(int));
  for (int t = 0; t < npix; t++)
    if (shifted[t] >= 0) counts_s[shifted[t]]++;
  int rc = orc_fill_missing(filled, centers, height, width, 1000);
  if (rc == 0) {
    orc_invalidate_disconnected(filled, height, width, nspix);
    rc = orc_fill_missing(filled, centers, height, width, 1000);
  }
  if (rc != 0) {
    free(fsum); free(counts); free(flow_ds); free(centers); free(shifted); free(counts_s);
    return rc;
  }
  int* counts_f = (int*)calloc((size_t)nspix, sizeof(int));
  for (int t = 0; t < npix; t++)
    if (filled[t] >= 0) counts_f[filled[t]]++;
  for (int k = 0; k < nspix; k++) {
    prev->invalid_counts[k] =
        (float)((double)(counts_f[k] - counts_s[k]) / ((double)(float)counts_f[k] + 1e-8));
    if (counts[k] > 0)
      prev->prior_counts[k] = (float)fmin(fmax(1.0 * counts[k], 100.0), 10000.0);
  }
  if (shifted_out) memcpy(shifted_out, shifted, sizeof(int) * (size_t)npix);
  if (filled_out) memcpy(filled_out, filled, sizeof(int) * (size_t)npix);
  /* adaptive iteration count, pyapi.cu:139-149 */
  int ninvalid = 0;
  for (int t = 0; t < npix; t++) ninvalid += shifted[t] == -1;
  float iperc = (float)(ninvalid / (1.0 * npix));
  if (iperc > 0.20f)
    *niters_state = 12;
  else if (iperc < 0.01f)
    *niters_state = 4;
  else
    *niters_state = 8;
  int niters = cfg->prop_niters > 0 ? cfg->prop_niters : *niters_state;

  /* run_prop, prop.cu:677-865 */
  int nbuf = nspix * 10;
  osp* P = (osp*)calloc((size_t)nbuf, sizeof(osp));
  ostat* st = (ostat*)calloc((size_t)nbuf, sizeof(ostat));
  unsigned char* border = (unsigned char*)calloc((size_t)npix, 1);
  sm_work w = sm_alloc(npix, nbuf);
  w.shape_marg = cfg->shape_marginal;
  sweep_buf sb = sweep_alloc(nbuf);
  /* get_vectors_as_params + write_params, sparams_io.cu:123-140, 352-501 */
  for (int ix = 0; ix < nspix; ix++) {
    osp* p = P + ix;
    for (int c = 0; c < 3; c++) {
      p->mu_app[c] = prev->mu_app[3 * ix + c];
      p->prior_mu_app[c] = prev->prior_mu_app[3 * ix + c];
      p->sigma_shape[c] = prev->sigma_shape[3 * ix + c];
      p->prior_sigma_shape[c] = prev->prior_sigma_shape[3 * ix + c];
      p->sample_sigma_shape[c] = prev->sample_sigma_shape[3 * ix + c];
    }
    for (int c = 0; c < 2; c++) {
      p->mu_shape[c] = prev->mu_shape[2 * ix + c];
      p->prior_mu_shape[c] = prev->prior_mu_shape[2 * ix + c];
    }
    p->logdet_sigma_shape = (double)prev->logdet_sigma_shape[ix];
    /* invert_cov, sparams_io.cu:504-520 (off-diagonal not negated) */
    double* cv = p->prior_sigma_shape;
    double pr_det = cv[0] * cv[2] - cv[1] * cv[1];
    p->prior_icov[0] = cv[2] / pr_det;
    p->prior_icov[1] = cv[1] / pr_det;
    p->prior_icov[2] = cv[0] / pr_det;
    p->count = prev->counts[ix];
    p->icount = prev->invalid_counts[ix];
    p->sm_count = prev->sm_counts[ix];
    p->prior_count = prev->prior_counts[ix];
    p->prop = 1;
    p->valid = 0;
  }
  /* mark_active, init_sparams.cu:195-286 */
  for (int k = 0; k < nbuf; k++) {
    P[k].valid = 0;
    P[k].prop = 0;
    if (k >= nspix) {
      P[k].prior_sigma_shape[0] = 1.0 / sp_size;
      P[k].prior_sigma_shape[1] = 0;
      P[k].prior_sigma_shape[2] = 1.0 / sp_size;
    }
  }
  for (int t = 0; t < npix; t++) {
    int L = filled[t];
    if (L >= 0 && L < nbuf) {
      P[L].valid = 1;
      P[L].prop = 1;
    }
  }
  /* prop_bass_v3, prop.cu:326-468 */
  nig_consts nc = make_nig(cfg->sigma2_app);
  double log_alpha = stsp_det_log((double)expf(cfg->alpha_hastings));
  int smc = 1, max_spix = nspix - 1;
  float split_alpha = cfg->split_alpha, merge_alpha = cfg->merge_alpha,
        thresh_new = cfg->thresh_new;
  const int controlled = cfg->target_nspix > 0; /* prop.cu:356-360 */
  const int og_niters = niters;
  if (controlled) niters = 5000;
  for (int idx = 0; idx < niters; idx++) {
    update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    if (controlled && idx % 2 == 0 &&
        control_nspix(count_living(seg, npix, nbuf), cfg->target_nspix, idx, og_niters, 1,
                      &split_alpha, &merge_alpha, &thresh_new))
      break;
    if (is_split_iter(idx, cfg)) {
      smc += 1; /* run_split_p, split_merge_prop.cu:31-59 */
      int direction = smc % 2 + 1;
      max_spix = split_step(img, seg, shifted, P, st, nbuf, max_spix, direction, &nc,
                            split_alpha, sp_size, 1, width, height, &w);
      update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    }
    if (is_merge_iter(idx, cfg)) {
      max_spix = relabel_step(seg, npix, P, nbuf, prev, cfg->thresh_relabel, thresh_new,
                              nspix_prev, max_spix);
      smc += 1; /* run_merge_p, split_merge_prop.cu:62-85 */
      int direction = smc % 2;
      merge_step(img, seg, border, P, st, nbuf, direction, &nc, log_alpha, merge_alpha,
                 1, width, height, &w);
      update_params(img, seg, npix, width, P, st, nbuf, sp_size);
    }
    update_seg_P(img, seg, border, P, nbuf, &sb, cfg->niters_seg, cfg->sigma2_app,
                 cfg->potts, height, width);
  }
  update_params(img, seg, npix, width, P, st, nbuf, sp_size);
  store_sample_sigma_shape(P, st, nbuf);
  int ml = max_label(seg, npix);
  nspix = compactify(seg, npix, P, nbuf, nspix_prev, ml > max_spix ? ml : max_spix);
  rc = nspix;
  if (nspix > cap) {
    rc = -2;
  } else {
    export_params(P, nspix, out);
    update_prior(out, seg, npix, nspix);
  }
  sweep_free(&sb);
  sm_free(&w);
  free(border);
  free(st);
  free(P);
  free(fsum);
  free(counts);
  free(flow_ds);
  free(centers);
  free(shifted);
  free(counts_s);
  free(counts_f);
  return rc;
}

/* ------------------------------------------------------------------ */
/* unit entry points: single steps of the restatement, for the tests   */
/* that pin it against closed forms and brute force                    */
/* ------------------------------------------------------------------ */
float orc_calc_joint(const float* imgC, int x, int y, const float* mu_app, const double* mu_shape,
                     const double* isig, double logdet, int count, float sigma2_app,
                     float neigh_neq, float potts) {
  sweep_params sp = {mu_app, mu_shape, isig, &logdet, &count};
  float best = -INFINITY;
  int lab = -1;
  calc_joint(imgC, x, y, &sp, 0, sigma2_app, neigh_neq, potts, &best, &lab);
  return best;
}

double orc_nig_marginal(float sigma2_app, double n, const double* bn) {
  nig_consts nc = make_nig(sigma2_app);
  return nig_marginal(&nc, n, bn);
}

void orc_calc_simple_update(const int64_t* stat, int prop, float prior_count,
                            const double* prior_icov, const double* sample_sigma_shape,
                            int sp_size, float* mu_app, double* mu_shape, double* isig,
                            double* logdet) {
  osp p;
  memset(&p, 0, sizeof(p));
  ostat st;
  memset(&st, 0, sizeof(st));
  st.count = (int)stat[0];
  for (int c = 0; c < 3; c++) st.S[c] = stat[1 + c];
  st.X = stat[4];
  st.Y = stat[5];
  st.XX = stat[6];
  st.XY = stat[7];
  st.YY = stat[8];
  p.count = st.count;
  p.valid = 1;
  p.prop = prop;
  p.prior_count = prior_count;
  memcpy(p.prior_icov, prior_icov, sizeof(double) * 3);
  memcpy(p.sample_sigma_shape, sample_sigma_shape, sizeof(double) * 3);
  calc_simple_update(&p, &st, sp_size);
  memcpy(mu_app, p.mu_app, sizeof(float) * 3);
  memcpy(mu_shape, p.mu_shape, sizeof(double) * 2);
  memcpy(isig, p.sigma_shape, sizeof(double) * 3);
  *logdet = p.logdet_sigma_shape;
}

void orc_invert_cov(const double* cov, double* out) {
  double pr_det = cov[0] * cov[2] - cov[1] * cov[1];
  out[0] = cov[2] / pr_det;
  out[1] = cov[1] / pr_det;
  out[2] = cov[0] / pr_det;
}

/* the per-label mean flow of orc_prop (run_sp_downsample, rule R1) */
void orc_flow_pool(const float* flow, const int* seg, int npix, int nspix, float* flow_ds,
                   int* counts) {
  int64_t* fsum = (int64_t*)calloc((size_t)2 * nspix, sizeof(int64_t));
  memset(counts, 0, sizeof(int) * (size_t)nspix);
  for (int t = 0; t < npix; t++) {
    int L = seg[t];
    if (L < 0 || L >= nspix) continue;
    fsum[2 * L] += (int64_t)llrint((double)flow[2 * t] * FX_FLOW);
    fsum[2 * L + 1] += (int64_t)llrint((double)flow[2 * t + 1] * FX_FLOW);
    counts[L]++;
  }
  for (int k = 0; k < nspix; k++)
    for (int c = 0; c < 2; c++) {
      double sum = (double)fsum[2 * k + c] * (1.0 / FX_FLOW);
      flow_ds[2 * k + c] = counts[k] > 0 ? (float)(sum / (1. * counts[k])) : 0.0f;
    }
  free(fsum);
}

/* per-label working state of labels 0..K-1 of seg, as update_params leaves it */
static osp* unit_state(const float* img, const int* seg, int npix, int width, int K,
                       const int* prop, int nbuf, ostat** st_out) {
  osp* P = (osp*)calloc((size_t)nbuf, sizeof(osp));
  ostat* st = (ostat*)calloc((size_t)nbuf, sizeof(ostat));
  stats_accumulate(img, seg, npix, width, st, nbuf, NULL);
  for (int k = 0; k < K; k++) {
    P[k].count = st[k].count;
    P[k].valid = st[k].count > 0;
    P[k].prop = prop ? prop[k] : 0;
    if (st[k].count > 0) {
      P[k].mu_shape[0] = (double)st[k].X / st[k].count;
      P[k].mu_shape[1] = (double)st[k].Y / st[k].count;
    }
  }
  *st_out = st;
  return P;
}

int orc_split_eval(const float* img, const int* seg, const int* shifted, int height, int width,
                   int K, const int* prop, int direction, float sigma2_app, float split_alpha,
                   int prop_mode, int* halves, double* hasting, int* accept, int* select) {
  int npix = width * height, nbuf = 2 * K + 2, max_spix = K - 1;
  ostat* st;
  osp* P = unit_state(img, seg, npix, width, K, prop, nbuf, &st);
  sm_work w = sm_alloc(npix, nbuf);
  nig_consts nc = make_nig(sigma2_app);
  sm_clear(&w, nbuf);
  split_halves(seg, P, nbuf, max_spix, direction, width, height, &w);
  sm_accumulate(img, w.halves, prop_mode ? shifted : NULL, npix, width, &w);
  int* acc = (int*)calloc((size_t)nbuf, sizeof(int));
  split_eval(P, st, nbuf, max_spix, &nc, split_alpha, prop_mode, &w, acc);
  memcpy(halves, w.halves, sizeof(int) * (size_t)npix);
  for (int k = 0; k < K; k++) {
    hasting[k] = w.sm[k].hasting;
    accept[k] = acc[k];
    select[k] = w.sm[k].select;
  }
  free(acc);
  sm_free(&w);
  free(st);
  free(P);
  return 0;
}

int orc_merge_eval(const float* img, const int* seg, int height, int width, int K,
                   const int* prop, int direction, float sigma2_app, float alpha_hastings,
                   float merge_alpha, int prop_mode, int* cand, double* hasting, int* winner) {
  int npix = width * height, nbuf = 2 * K + 2;
  ostat* st;
  osp* P = unit_state(img, seg, npix, width, K, prop, nbuf, &st);
  sm_work w = sm_alloc(npix, nbuf);
  nig_consts nc = make_nig(sigma2_app);
  unsigned char* border = (unsigned char*)malloc((size_t)npix);
  find_border_pixels(seg, border, height, width);
  int* win = (int*)malloc(sizeof(int) * (size_t)nbuf);
  merge_eval(img, seg, border, P, st, nbuf, direction, &nc,
             stsp_det_log((double)expf(alpha_hastings)), merge_alpha, prop_mode, width, height,
             &w, win);
  for (int k = 0; k < K; k++) {
    cand[k] = w.pairs[2 * k + 1];
    hasting[k] = w.sm[k].hasting;
    winner[k] = win[k];
  }
  free(win);
  free(border);
  sm_free(&w);
  free(st);
  free(P);
  return 0;
}

int orc_relabel_search(const float* mu_app, const unsigned char* living, int max_spix,
                       const float* prior_mu_app, int nspix_prev, float thresh_relabel,
                       float thresh_new, int nbuf, int* target) {
  return relabel_search(mu_app, 3, living, max_spix, prior_mu_app, nspix_prev, thresh_relabel,
                        thresh_new, nbuf, target);
}
