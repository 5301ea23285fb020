// The per-step label ops of the st_spix.prop pipeline (its prop_cuda module),
// as standalone entry points: shift_labels (csrc/prop/shift_labels.cu) and
// fill_missing (csrc/prop/fill_missing.cu).  The fused BIST propagation
// (host.hip, k_prop.hip) has its own shift and fill; these follow the
// prop_cuda forms, which differ from them (integer flow per label with
// collision counts; float centres over a list of missing pixels).
#include <algorithm>

#include "../../include/stspix.h"
#include "common.h"

namespace stsp {

// ---------------------------------------------------------------------------
// shift_labels (shift_labels.cu:28-62): a pixel of label L moves by L's
// integer flow; the destination counts its arrivals and keeps the largest
// label.  Both are order-free (integer add and max), so the result does not
// depend on the schedule.
// ---------------------------------------------------------------------------
__global__ void ks_shift_init(int* __restrict__ shifted, int* __restrict__ counts, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  shifted[i] = -1;
  counts[i] = 0;
}

__global__ void ks_shift_labels(const int* __restrict__ spix, const int* __restrict__ flow,
                                int nspix, int height, int width, int* __restrict__ shifted,
                                int* __restrict__ counts) {
  const int npix = height * width;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= npix) return;
  const int L = spix[(size_t)b * npix + i];
  if (L < 0 || L >= nspix) return;  // no flow entry: the pixel does not move
  const int* fl = flow + 2 * ((size_t)b * nspix + L);
  const int hd = i / width + fl[1], wd = i % width + fl[0];
  if (hd < 0 || hd >= height || wd < 0 || wd >= width) return;
  const size_t d = (size_t)b * npix + (size_t)hd * width + wd;
  atomicAdd(counts + d, 1);
  atomicMax(shifted + d, L);
}

// ---------------------------------------------------------------------------
// fill_missing (fill_missing.cu:49-110): rounds of a border pass over the
// listed pixels (find_border_along_missing, 276-314) and four checkerboard
// passes (update_missing_seg_nn, 142-273), until a round counts no -1 among
// the listed pixels, counts as many as the round before (the first round
// compares with 1), or break_iter rounds ran.  The rounds run in chunks;
// each round's kernels read the stop state on the device, so a chunk can
// overrun the stop without effect, and the host reads the state once per
// chunk.
// ---------------------------------------------------------------------------
struct FillState {
  int alive;  // the next round runs
  int prev;   // the count the next round compares with
  int rounds; // rounds run
  int pad;
};
constexpr int kFillSub = 64;  // partial -1 counters per round
constexpr int kFillChunk = 8;

__global__ void ks_fill_start(FillState* st, int* part) {
  const int t = threadIdx.x;
  if (t == 0) *st = FillState{1, 1, 0, 0};
  for (int i = t; i < kFillSub * kFillChunk; i += blockDim.x) part[i] = 0;
}

__global__ void ks_fill_border(const int* __restrict__ seg, const int* __restrict__ missing,
                               int nmissing, int height, int width, uint8_t* __restrict__ border,
                               const FillState* __restrict__ st, int* __restrict__ part) {
  if (!st->alive) return;  // grid-uniform
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  bool neg = false;
  const int idx = j < nmissing ? missing[j] : -1;
  if (idx >= 0 && idx < height * width) {  // entries outside the image are ignored
    const int x = idx % width, y = idx / width;
    const int C = seg[idx];
    const int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
    const int S = y < height - 1 ? seg[idx + width] : -1;
    const int E = x < width - 1 ? seg[idx + 1] : -1;
    const bool valid = N >= 0 || Wl >= 0 || S >= 0 || E >= 0;
    border[idx] = (valid && C < 0) ? 1 : 0;  // listed pixels only: the rest stays 0
    neg = C < 0;
  }
  // one atomic per wave, spread over kFillSub counters
  const unsigned long long m = __ballot(neg);
  if ((threadIdx.x & 63) == 0 && m)
    atomicAdd(part + ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kFillSub - 1)),
              __popcll(m));
}

__global__ void ks_fill_phase(int* __restrict__ seg, const float* __restrict__ centers,
                              int ncenters, const int* __restrict__ missing, int nmissing,
                              int height, int width, const uint8_t* __restrict__ border, int xm,
                              int ym, const FillState* __restrict__ st) {
  if (!st->alive) return;  // grid-uniform
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nmissing) return;
  const int idx = missing[j];
  if (idx < 0 || idx >= height * width) return;
  const int x = idx % width, y = idx / width;
  if ((x & 1) != xm || (y & 1) != ym || !border[idx]) return;
  // the 4-neighbours have the other parity: no pass of this phase writes them
  const int N = y > 0 ? seg[idx - width] : -1, Wl = x > 0 ? seg[idx - 1] : -1;
  const int E = x < width - 1 ? seg[idx + 1] : -1, S = y < height - 1 ? seg[idx + width] : -1;
  const int cand[4] = {N, S, Wl, E};
  float best = -999999.0f, lab = -1.0f;  // the reference keeps the label in a float
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int L = cand[c];
    bool dup = false;
#pragma unroll
    for (int q = 0; q < c; q++) dup = dup || cand[q] == L;
    if (L < 0 || dup || L >= ncenters) continue;
    // isotropic_space (fill_missing.cu:127-140), float arithmetic
    const float dx = (1.0f * (float)x - centers[2 * L]) / (1.0f * (float)width);
    const float dy = (1.0f * (float)y - centers[2 * L + 1]) / (1.0f * (float)height);
    const float sim = -dx * dx - dy * dy;
    if (sim > best) {
      best = sim;
      lab = (float)L;
    }
  }
  seg[idx] = (int)lab;
}

// end of round r of the chunk: the stop rule of the host loop
__global__ void ks_fill_end(FillState* st, int* part, int r) {
  if (threadIdx.x != 0 || !st->alive) return;
  int n = 0;
  for (int i = 0; i < kFillSub; i++) n += part[r * kFillSub + i];
  st->rounds += 1;
  if (n == 0 || n == st->prev) st->alive = 0;
  st->prev = n;
}

__global__ void ks_fill_left(const int* __restrict__ seg, const int* __restrict__ missing,
                             int nmissing, int npix, int* left) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int idx = j < nmissing ? missing[j] : -1;
  const bool neg = idx >= 0 && idx < npix && seg[idx] < 0;
  const unsigned long long m = __ballot(neg);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(left, __popcll(m));
}

}  // namespace stsp

using namespace stsp;

extern "C" {

int stsp_shift_labels(const int* spix, const int* flow, int nbatch, int height, int width,
                      int nspix, int* shifted, int* counts, hipStream_t s) {
  if (!spix || !flow || !shifted || !counts || nbatch < 1 || height < 1 || width < 1 ||
      nspix < 0 || (long long)height * width > (1ll << 31) - 1 || nbatch > 65535)
    return STSP_ERR_ARG;
  const long long n = (long long)nbatch * height * width;
  ks_shift_init<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(shifted, counts, n);
  const dim3 grid((unsigned)(((long long)height * width + kBlock - 1) / kBlock), (unsigned)nbatch);
  ks_shift_labels<<<grid, kBlock, 0, s>>>(spix, flow, nspix, height, width, shifted, counts);
  return hipGetLastError() == hipSuccess ? STSP_OK : STSP_ERR_HIP;
}

int stsp_fill_missing(int* seg, const float* centers, int ncenters, const int* missing,
                      int nmissing, int height, int width, int break_iter, int* rounds_out,
                      int* negatives_out, hipStream_t s) {
  if (!seg || (!centers && ncenters > 0) || (!missing && nmissing > 0) || nmissing < 0 ||
      ncenters < 0 || height < 1 || width < 1 || (long long)height * width > (1ll << 31) - 1)
    return STSP_ERR_ARG;
  if (rounds_out) *rounds_out = 0;
  if (negatives_out) *negatives_out = 0;
  if (nmissing == 0) return STSP_OK;  // the reference runs no round then
  const size_t npix = (size_t)height * width;
  // workspace: border flags, stop state, partial counts, host mirror
  const size_t wbytes = 256 + npix + sizeof(int) * kFillSub * kFillChunk + 256;
  char* w = nullptr;
  if (hipMallocAsync((void**)&w, wbytes, s) != hipSuccess) return STSP_ERR_HIP;
  FillState* st = reinterpret_cast<FillState*>(w);
  int* left = reinterpret_cast<int*>(w + 128);
  int* part = reinterpret_cast<int*>(w + 256);
  uint8_t* border = reinterpret_cast<uint8_t*>(w + 256 + sizeof(int) * kFillSub * kFillChunk);
  int rc = STSP_OK;
  const unsigned nb = (unsigned)((nmissing + kBlock - 1) / kBlock);
  ks_fill_start<<<1, kBlock, 0, s>>>(st, part);
  (void)hipMemsetAsync(border, 0, npix, s);
  FillState host = {1, 1, 0, 0};
  int launched = 0;
  while (host.alive && (break_iter <= 0 || launched < break_iter)) {
    const int nr = break_iter > 0 ? std::min(kFillChunk, break_iter - launched) : kFillChunk;
    (void)hipMemsetAsync(part, 0, sizeof(int) * kFillSub * kFillChunk, s);
    for (int r = 0; r < nr; r++) {
      ks_fill_border<<<nb, kBlock, 0, s>>>(seg, missing, nmissing, height, width, border, st,
                                            part + r * kFillSub);
      for (int xm = 0; xm < 2; xm++)
        for (int ym = 0; ym < 2; ym++)
          ks_fill_phase<<<nb, kBlock, 0, s>>>(seg, centers, ncenters, missing, nmissing, height,
                                               width, border, xm, ym, st);
      ks_fill_end<<<1, 64, 0, s>>>(st, part, r);
    }
    launched += nr;
    if (hipMemcpyAsync(&host, st, sizeof(FillState), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = STSP_ERR_HIP;
      break;
    }
  }
  int nleft = 0;
  if (rc == STSP_OK) {
    (void)hipMemsetAsync(left, 0, sizeof(int), s);
    ks_fill_left<<<nb, kBlock, 0, s>>>(seg, missing, nmissing, (int)npix, left);
    if (hipMemcpyAsync(&nleft, left, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = STSP_ERR_HIP;
  }
  (void)hipFreeAsync(w, s);
  if (rc == STSP_OK && hipGetLastError() != hipSuccess) rc = STSP_ERR_HIP;
  if (rounds_out) *rounds_out = host.rounds;
  if (negatives_out) *negatives_out = nleft;
  return rc;
}

}  // extern "C"
