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
void launch_bass_mark(Params P, int nspix, int nbuf, int sp_size, hipStream_t s);
void launch_bass_init(Params P, int nspix, int nbuf, int sp_size, int npix, hipStream_t s);
void launch_scan(const int* in, int* out, int n, int* total, hipStream_t s);
void launch_fill(int* p, int v, int n, hipStream_t s);
void launch_mark_alive(const int* seg, int npix, int* alive, int n, hipStream_t s);
void launch_compact(int* seg, int npix, Params* P, int n, int prev, int nbuf, int* alive,
                    int* flags, int* rank, int* num_new, void* tmp, hipStream_t s);
size_t compact_tmp_bytes(int n);
void launch_export(Params P, OutParams o, const int* nspix, const int* alive, int nmax,
                   hipStream_t s);

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
// colour preprocessing (k_color.hip)
void launch_rescale_u8(const uint8_t* in, float* out, long long n, float scale, hipStream_t s);
void launch_rgb_to_lab(const float* bgr, float* lab, long long npix, hipStream_t s);
void launch_lab_to_rgb(const float* lab, float* rgb, long long npix, hipStream_t s);

}  // namespace stsp

namespace stsp {
void launch_marked(const float* vid, const int* spix, const float* color, int T, int H, int W,
                   float* marked, hipStream_t s);
}

