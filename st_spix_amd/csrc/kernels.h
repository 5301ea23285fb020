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

// the compaction and export of every frame of a batch (one launch per stage);
// frame f owns label blocks [off_lab[f], off_lab[f + 1]) and pixel blocks
// [off_pix[f], off_pix[f + 1]) of kBlock threads
struct EpiFrame {
  Params P;
  OutParams o;
  int* seg;       // final labels, compacted in place
  int* spix_out;  // the caller's label map
  int *alive, *flags, *rank, *num_new, *total;
  void* ctmp;     // compact_tmp_bytes(nbuf)
  int npix, nbuf, prev;
};
struct EpiBatch {
  const EpiFrame* fr;
  const int* off_lab;
  const int* off_pix;
  int* counts;  // per frame: the number of new labels
  int nf, lab_blocks, pix_blocks;
};
void epi_launch(const EpiBatch& e, hipStream_t s);

// k_prop.hip: the propagation prologue of every frame of a batch, one
// launch per stage.  Each frame's pointers and sizes are one ProFrame entry
// of a device table; ProOff holds, per block kind, the first block of each
// frame (prefix sums, kProMax + 1 entries each).
constexpr int kProMax = 64;     // frames per batch
constexpr int kFillMax = 1000;  // hole-filling rounds (fill_missing.cu break_iter)
struct ProFrame {
  int W, H, npix, np, nbuf, sp_size;
  const float* flow;       // [npix, 2]
  const int* spix_prev;    // [npix]
  const double* mu_shape_prev;
  float* prior_counts;     // previous parameters, updated in place
  float* invalid_counts;
  InParams in;             // previous parameters
  Params P;                // working parameters (nbuf slots)
  int64_t* fsum;
  int *counts_prev, *counts_s, *counts_f;
  float* flow_ds;
  double* centers;
  unsigned long long* keys;
  int *shifted, *seg, *parent, *rsize, *max_spix;
  uint8_t* border;
  unsigned long long* best;
};
enum { kProPix = 0, kProLab, kProBuf, kProTile, kProKinds };
// A frame's holes are kept in kHoleSub sub-lists (pixel block b appends to
// sub-list b % kHoleSub, so the appends of a frame spread over kHoleSub
// counters).
constexpr int kHoleSub = 64;
// partial hole counters per frame: the counting blocks of a frame spread their
// atomics over kNegSub addresses instead of one
constexpr int kNegSub = 64;

struct Pro {
  const ProFrame* fr;  // device table
  const int* off;      // [kProKinds][kProMax + 1] over the frames launched
  int nf;              // frames launched (entries of off)
  int nall;            // frames in the table
  int* fillres;        // [nall][2]: filling rounds run, 1 if the hole count stalled
  int* negs;           // [nall * kNegSub]: holes of the shifted map, kNegSub partial
                       // counters per frame (summed by the host)
  int* hlen;           // [nall * kHoleSub]: lengths of each frame's hole sub-lists
                       // (pixel indices of the holes when a fill starts, kept in
                       // the frame's `keys`)
};
// host side: block counts per kind for a frame
int pro_blocks(int kind, int W, int H, int np, int nbuf);
// the whole prologue except the hole-filling rounds
void pro_shift(const Pro& p, const int* total, hipStream_t s);
void pro_hole_list(const Pro& p, const int* total, hipStream_t s);
void pro_fill_all(const Pro& p, hipStream_t s);
void pro_disconnected(const Pro& p, const int* total, hipStream_t s);
void pro_finish(const Pro& p, const int* total, hipStream_t s);
// colour preprocessing (k_color.hip)
void launch_rescale_u8(const uint8_t* in, float* out, long long n, float scale, hipStream_t s);
void launch_rgb_to_lab(const float* bgr, float* lab, long long npix, hipStream_t s);
void launch_lab_to_rgb(const float* lab, float* rgb, long long npix, hipStream_t s);

}  // namespace stsp

namespace stsp {
void launch_marked(const float* vid, const int* spix, const float* color, int T, int H, int W,
                   float* marked, hipStream_t s);
}

