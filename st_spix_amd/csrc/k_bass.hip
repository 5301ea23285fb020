// Per-frame kernels around the refinement loop: hexagonal initialisation,
// first-frame parameter initialisation, scans, compaction, parameter export
// and the border overlay.  The refinement loop itself is in k_batch.hip.
#include "kernels.h"

namespace stsp {

// ---------------------------------------------------------------------------
// hexagonal initialisation (reference init_seg.cu:20-81)
// The reference compares each pixel with every centre (O(npix*K)); the
// nearest centre always lies in the 5x5 block of hex cells around the pixel,
// so only those 25 are compared, with the same distance expression and the
// same tie rule (last index wins).
// ---------------------------------------------------------------------------
__global__ void k_hex_init(int* seg, int H, int W, int nx, int ny, double Hh, double w) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= H * W) return;
  int x = idx % W, y = idx / W;
  int r0 = (int)floor((double)y / (1.5 * Hh));
  double D2 = 1.79769313486231570815e+308;
  int best = -1;
  for (int r = r0 - 2; r <= r0 + 2; r++) {
    if (r < 0 || r >= ny) continue;
    double off = (r % 2 == 0) ? 0.5 * w : 0.0;
    int c0 = (int)floor(((double)x - off) / w);
    for (int c = c0 - 2; c <= c0 + 2; c++) {
      if (c < 0 || c >= nx) continue;
      int j = r * nx + c;
      double cx = (double)c * w;
      double cy = (double)r * 1.5 * Hh;
      if (r % 2 == 0) cx = cx + 0.5 * w;
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

void launch_hex_init(int* seg, int H, int W, int sp_size, hipStream_t s) {
  HexGeom g = hex_geometry(sp_size, W, H);
  k_hex_init<<<div_up((long long)H * W, kBlock), kBlock, 0, s>>>(seg, H, W, g.nx, g.ny, g.Hh,
                                                                 g.w);
}

HexGeom hex_geometry(int sp_size, int width, int height) {
  HexGeom g;
  g.Hh = sqrt((double)pow((double)sp_size, 2) / (1.5 * sqrt(3.0)));
  g.w = sqrt(3.0) * g.Hh;
  g.nx = (int)floor((double)width / g.w) + 2;
  g.ny = (int)floor((double)height / (1.5 * g.Hh)) + 2;
  return g;
}

// ---------------------------------------------------------------------------
// first-frame parameter initialisation (init_sparams.cu:25-166, 195-286)
// ---------------------------------------------------------------------------
__global__ void k_bass_mark(Params P, int nspix, int nbuf, int sp_size) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  // zero everything the reference leaves to the allocator (rule R11)
  P.icount[k] = 0.f;
  P.sm_count[k] = 0;
  for (int c = 0; c < 3; c++) {
    P.prior_mu_app[3 * k + c] = 0.f;
    P.sample_sigma_shape[3 * k + c] = 0.0;
    P.prior_icov[3 * k + c] = 0.0;
  }
  P.prior_mu_shape[2 * k] = P.prior_mu_shape[2 * k + 1] = 0.0;
  P.logdet[k] = 0.0;
  P.prior_count[k] = 0.f;
  P.valid[k] = k < nspix ? 1 : 0;
  P.prop[k] = k < nspix ? 1 : 0;
  if (k >= nspix) {
    P.prior_sigma_shape[3 * k] = 1.0 / sp_size;
    P.prior_sigma_shape[3 * k + 1] = 0;
    P.prior_sigma_shape[3 * k + 2] = 1.0 / sp_size;
  } else {
    P.prior_sigma_shape[3 * k] = P.prior_sigma_shape[3 * k + 1] = P.prior_sigma_shape[3 * k + 2] = 0;
  }
}

__global__ void k_bass_init(Params P, int nspix, int nbuf, int sp_size, int npix) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbuf) return;
  int count = (int)(npix / (1. * nspix));
  P.prior_sigma_shape[3 * k] = 1.0 / sp_size;
  P.prior_sigma_shape[3 * k + 1] = 0.;
  P.prior_sigma_shape[3 * k + 2] = 1.0 / sp_size;
  if (k < nspix) {
    P.valid[k] = 1;
    P.prior_count[k] = (float)count;
    for (int c = 0; c < 3; c++) {
      P.prior_mu_app[3 * k + c] = P.mu_app[3 * k + c];
      P.mu_app[3 * k + c] = 0.f;
    }
    P.prior_icov[3 * k] = 1;
    P.prior_icov[3 * k + 1] = 0;
    P.prior_icov[3 * k + 2] = 1;
    P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0;
    P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0;
    P.prop[k] = 0;
  } else {
    P.count[k] = 0;
    for (int c = 0; c < 3; c++) {
      P.mu_app[3 * k + c] = 0.f;
      P.prior_mu_app[3 * k + c] = 0.f;
    }
    P.mu_shape[2 * k] = P.mu_shape[2 * k + 1] = 0;
    P.prior_mu_shape[2 * k] = P.prior_mu_shape[2 * k + 1] = 0;
    P.sigma_shape[3 * k] = P.sigma_shape[3 * k + 1] = P.sigma_shape[3 * k + 2] = 0;
    P.logdet[k] = 0;
    P.prop[k] = 0;
    P.valid[k] = 0;
    P.prior_icov[3 * k] = 1;
    P.prior_icov[3 * k + 1] = 0;
    P.prior_icov[3 * k + 2] = 1;
  }
}

void launch_bass_mark(Params P, int nspix, int nbuf, int sp_size, hipStream_t s) {
  k_bass_mark<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, nspix, nbuf, sp_size);
}

void launch_bass_init(Params P, int nspix, int nbuf, int sp_size, int npix, hipStream_t s) {
  k_bass_init<<<div_up(nbuf, kBlock), kBlock, 0, s>>>(P, nspix, nbuf, sp_size, npix);
}

// ---------------------------------------------------------------------------
// scans and compaction (compact_spix.cu:41-122, 196-295)
// ---------------------------------------------------------------------------
// Single-workgroup exclusive scan of n ints (n up to a few million); writes
// the total to *total if non-null.
__global__ void __launch_bounds__(1024) k_scan1(const int* __restrict__ in, int* __restrict__ out,
                                                int n, int* total) {
  __shared__ int part[1024];
  int tid = threadIdx.x;
  int chunk = (n + 1023) / 1024;
  int b = tid * chunk, e = min(n, b + chunk);
  int s = 0;
  for (int i = b; i < e; i++) s += in[i];
  part[tid] = s;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    int v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - s;
  for (int i = b; i < e; i++) {
    int v = in[i];
    out[i] = run;
    run += v;
  }
  if (tid == 1023 && total) *total = part[1023];
}

void launch_scan(const int* in, int* out, int n, int* total, hipStream_t s) {
  k_scan1<<<1, 1024, 0, s>>>(in, out, n, total);
}

__global__ void k_fill_i32(int* p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill(int* p, int v, int n, hipStream_t s) {
  if (n > 0) k_fill_i32<<<div_up(n, kBlock), kBlock, 0, s>>>(p, v, n);
}

__global__ void k_mark_alive(const int* __restrict__ seg, int npix, int* alive, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int k = seg[t];
  if (k >= 0 && k < n) alive[k] = 1;
}
void launch_mark_alive(const int* seg, int npix, int* alive, int n, hipStream_t s) {
  k_mark_alive<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, npix, alive, n);
}

// alive flags of ids >= prev become "new" flags; ids below prev are kept
__global__ void k_new_flags(const int* alive, int* flags, int n, int prev) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  flags[k] = (k >= prev && alive[k]) ? 1 : 0;
}

__global__ void k_remap(int* seg, int npix, const int* rank, int prev) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npix) return;
  int k = seg[t];
  if (k >= prev) seg[t] = prev + rank[k];
}

// the parameter fields fill_new_params_from_old copies (compact_spix.cu:220-295)
struct CompactTmp {
  float mu_app[3], prior_mu_app[3];
  double mu_shape[2], sigma_shape[3], prior_mu_shape[2], prior_sigma_shape[3];
  double sample_sigma_shape[3], logdet;
  int count, valid;
  float prior_count;
};

__device__ __forceinline__ void compact_gather_one(const Params& P, int k, CompactTmp* dst) {
  CompactTmp t;
  for (int c = 0; c < 3; c++) {
    t.mu_app[c] = P.mu_app[3 * k + c];
    t.prior_mu_app[c] = P.prior_mu_app[3 * k + c];
    t.sigma_shape[c] = P.sigma_shape[3 * k + c];
    t.prior_sigma_shape[c] = P.prior_sigma_shape[3 * k + c];
    t.sample_sigma_shape[c] = P.sample_sigma_shape[3 * k + c];
  }
  for (int c = 0; c < 2; c++) {
    t.mu_shape[c] = P.mu_shape[2 * k + c];
    t.prior_mu_shape[c] = P.prior_mu_shape[2 * k + c];
  }
  t.logdet = P.logdet[k];
  t.count = P.count[k];
  t.valid = P.valid[k];
  t.prior_count = P.prior_count[k];
  *dst = t;
}

__device__ __forceinline__ void compact_scatter_one(const Params& P, int k, const CompactTmp& t) {
  for (int c = 0; c < 3; c++) {
    P.mu_app[3 * k + c] = t.mu_app[c];
    P.prior_mu_app[3 * k + c] = t.prior_mu_app[c];
    P.sigma_shape[3 * k + c] = t.sigma_shape[c];
    P.prior_sigma_shape[3 * k + c] = t.prior_sigma_shape[c];
    P.sample_sigma_shape[3 * k + c] = t.sample_sigma_shape[c];
  }
  for (int c = 0; c < 2; c++) {
    P.mu_shape[2 * k + c] = t.mu_shape[c];
    P.prior_mu_shape[2 * k + c] = t.prior_mu_shape[c];
  }
  P.logdet[k] = t.logdet;
  P.count[k] = t.count;
  P.valid[k] = t.valid;
  P.prior_count[k] = t.prior_count;
}

__global__ void k_compact_gather(Params P, const int* flags, const int* rank, int n, int prev,
                                 CompactTmp* tmp) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !flags[k]) return;
  compact_gather_one(P, k, tmp + rank[k]);
}

__global__ void k_compact_scatter(Params P, const CompactTmp* tmp, const int* num_new, int prev,
                                  int nbuf) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= *num_new || prev + d >= nbuf) return;
  compact_scatter_one(P, prev + d, tmp[d]);
}

void launch_compact(int* seg, int npix, Params* P, int n, int prev, int nbuf, int* alive,
                    int* flags, int* rank, int* num_new, void* tmp, hipStream_t s) {
  launch_fill(alive, 0, n, s);
  launch_mark_alive(seg, npix, alive, n, s);
  k_new_flags<<<div_up(n, kBlock), kBlock, 0, s>>>(alive, flags, n, prev);
  launch_scan(flags, rank, n, num_new, s);
  if (P) {
    k_compact_gather<<<div_up(n, kBlock), kBlock, 0, s>>>(*P, flags, rank, n, prev,
                                                          (CompactTmp*)tmp);
  }
  k_remap<<<div_up(npix, kBlock), kBlock, 0, s>>>(seg, npix, rank, prev);
  if (P) {
    k_compact_scatter<<<div_up(n, kBlock), kBlock, 0, s>>>(*P, (const CompactTmp*)tmp, num_new,
                                                           prev, nbuf);
  }
}

size_t compact_tmp_bytes(int n) { return sizeof(CompactTmp) * (size_t)(n > 0 ? n : 1); }

// ---------------------------------------------------------------------------
// parameter export (read_params, sparams_io.cu:189-286) and update_prior with
// invert=false (sparams_io.cu:583-639, rule R10)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void export_one(const Params& P, const OutParams& o, int ix, bool alive) {
  bool upd = alive && P.count[ix] > 0;
  for (int c = 0; c < 3; c++) {
    o.mu_app[3 * ix + c] = P.mu_app[3 * ix + c];
    o.prior_mu_app[3 * ix + c] = upd ? P.mu_app[3 * ix + c] : P.prior_mu_app[3 * ix + c];
    o.sigma_shape[3 * ix + c] = P.sigma_shape[3 * ix + c];
    o.prior_sigma_shape[3 * ix + c] =
        upd ? P.sigma_shape[3 * ix + c] : P.prior_sigma_shape[3 * ix + c];
    o.sample_sigma_shape[3 * ix + c] = P.sample_sigma_shape[3 * ix + c];
  }
  for (int c = 0; c < 2; c++) {
    o.mu_shape[2 * ix + c] = P.mu_shape[2 * ix + c];
    o.prior_mu_shape[2 * ix + c] = upd ? P.mu_shape[2 * ix + c] : P.prior_mu_shape[2 * ix + c];
  }
  o.logdet_sigma_shape[ix] = (float)P.logdet[ix];
  o.counts[ix] = P.count[ix];
  o.invalid_counts[ix] = P.icount[ix];
  o.sm_counts[ix] = P.sm_count[ix];
  o.prior_counts[ix] = P.prior_count[ix];
  o.ids[ix] = ix;
}

__global__ void k_export(Params P, OutParams o, const int* nspix_ptr, const int* alive) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x;
  int nspix = *nspix_ptr;
  if (ix >= nspix || ix >= o.cap) return;
  export_one(P, o, ix, alive[ix] != 0);
}

void launch_export(Params P, OutParams o, const int* nspix, const int* alive, int nmax,
                   hipStream_t s) {
  k_export<<<div_up(nmax, kBlock), kBlock, 0, s>>>(P, o, nspix, alive);
}

// ---------------------------------------------------------------------------
// The compaction and export of every frame of a batch, one launch per stage
// over all frames (the per-frame launch_compact, count, launch_mark_alive,
// launch_export and label copy, in the same order).  After the remap the
// live flags need no second pass over the labels: a label below prev keeps
// its pixels, and the new labels are exactly prev .. prev + num_new - 1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int epi_frame(const int* __restrict__ off, int nf) {
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
#define EPI_LAB                                                     \
  const int f = epi_frame(e.off_lab, e.nf);                         \
  const EpiFrame& E = e.fr[f];                                      \
  const int k = (blockIdx.x - e.off_lab[f]) * kBlock + threadIdx.x; \
  if (k >= E.nbuf) return;
#define EPI_PIX                                                     \
  const int f = epi_frame(e.off_pix, e.nf);                         \
  const EpiFrame& E = e.fr[f];                                      \
  const int t = (blockIdx.x - e.off_pix[f]) * kBlock + threadIdx.x; \
  if (t >= E.npix) return;

__global__ void ke_clear(EpiBatch e) {
  EPI_LAB
  E.alive[k] = 0;
}
__global__ void ke_mark(EpiBatch e) {
  EPI_PIX
  const int v = E.seg[t];
  if (v >= 0 && v < E.nbuf) E.alive[v] = 1;
}
__global__ void ke_flags(EpiBatch e) {
  EPI_LAB
  E.flags[k] = (k >= E.prev && E.alive[k]) ? 1 : 0;
}
// exclusive scan of each frame's flags (one workgroup per frame, as k_scan1)
__global__ void __launch_bounds__(1024) ke_scan(EpiBatch e) {
  __shared__ int part[1024];
  const EpiFrame& E = e.fr[blockIdx.x];
  const int n = E.nbuf, tid = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int b = tid * chunk, en = min(n, b + chunk);
  int sum = 0;
  for (int i = b; i < en; i++) sum += E.flags[i];
  part[tid] = sum;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const int v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;
  for (int i = b; i < en; i++) {
    const int v = E.flags[i];
    E.rank[i] = run;
    run += v;
  }
  if (tid == 1023) {
    const int tot = part[1023];
    *E.num_new = tot;
    *E.total = E.prev + tot;
    e.counts[blockIdx.x] = tot;
  }
}
__global__ void ke_gather(EpiBatch e) {
  EPI_LAB
  if (E.flags[k]) compact_gather_one(E.P, k, (CompactTmp*)E.ctmp + E.rank[k]);
}
// the remapped labels, also into the caller's label map
__global__ void ke_remap(EpiBatch e) {
  EPI_PIX
  int v = E.seg[t];
  if (v >= E.prev) {
    v = E.prev + E.rank[v];
    E.seg[t] = v;
  }
  E.spix_out[t] = v;
}
__global__ void ke_scatter(EpiBatch e) {
  EPI_LAB
  const int n_new = *E.num_new;
  if (k < n_new && E.prev + k < E.nbuf) compact_scatter_one(E.P, E.prev + k, ((const CompactTmp*)E.ctmp)[k]);
  if (k >= E.prev) E.alive[k] = k - E.prev < n_new ? 1 : 0;
}
__global__ void ke_export(EpiBatch e) {
  EPI_LAB
  if (k >= *E.total || k >= E.o.cap) return;
  export_one(E.P, E.o, k, E.alive[k] != 0);
}
#undef EPI_LAB
#undef EPI_PIX

void epi_launch(const EpiBatch& e, hipStream_t s) {
  ke_clear<<<e.lab_blocks, kBlock, 0, s>>>(e);
  ke_mark<<<e.pix_blocks, kBlock, 0, s>>>(e);
  ke_flags<<<e.lab_blocks, kBlock, 0, s>>>(e);
  ke_scan<<<e.nf, 1024, 0, s>>>(e);
  ke_gather<<<e.lab_blocks, kBlock, 0, s>>>(e);
  ke_remap<<<e.pix_blocks, kBlock, 0, s>>>(e);
  ke_scatter<<<e.lab_blocks, kBlock, 0, s>>>(e);
  ke_export<<<e.lab_blocks, kBlock, 0, s>>>(e);
}

}  // namespace stsp

namespace stsp {

// ---------------------------------------------------------------------------
// border overlay of get_marked_video (pyapi.cu:221-303) with
// find_border_pixels_end (seg_utils.cu:218-257)
// ---------------------------------------------------------------------------
__global__ void k_marked(const float* __restrict__ vid, const int* __restrict__ spix,
                         float c0, float c1, float c2, int H, int W, long long total,
                         float* __restrict__ marked) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int npix = H * W;
  int idx = (int)(t % npix);
  long long base = t - idx;
  int x = idx % W, y = idx / W;
  int C = spix[t];
  int N = y > 0 ? spix[base + idx - W] : -1;
  int Wl = x > 0 ? spix[base + idx - 1] : -1;
  int S = y < H - 1 ? spix[base + idx + W] : -1;
  int E = x < W - 1 ? spix[base + idx + 1] : -1;
  bool b = (N >= 0 && C > N) || (S >= 0 && C > S) || (E >= 0 && C > E) || (Wl >= 0 && C > Wl);
  if (b) {
    marked[3 * t] = c0;
    marked[3 * t + 1] = c1;
    marked[3 * t + 2] = c2;
  } else {
    for (int c = 0; c < 3; c++) marked[3 * t + c] = fminf(fmaxf(vid[3 * t + c], 0.0f), 1.0f);
  }
}

void launch_marked(const float* vid, const int* spix, const float* color, int T, int H, int W,
                   float* marked, hipStream_t s) {
  long long total = (long long)T * H * W;
  k_marked<<<(int)((total + kBlock - 1) / kBlock), kBlock, 0, s>>>(vid, spix, color[0], color[1],
                                                                   color[2], H, W, total, marked);
}

}  // namespace stsp
