/*
 * C ABI of the MI355X BASS/BIST superpixel library (libstspix.so).
 *
 * All array arguments are device pointers (HBM), row-major, channels-last:
 *   image  float32 [H, W, 3]
 *   flow   float32 [H, W, 2]   (u = dx, v = dy; flow of frame t-1 -> t)
 *   labels int32   [H, W]
 * Every call is asynchronous on the given hipStream_t except where a host
 * value is returned (nspix_out, niters_state), in which case the call waits
 * for the stream.  Calls on one context are not thread-safe; distinct
 * contexts may be used from distinct threads.  No C++ exception crosses the
 * boundary: every entry point returns an stsp_status code.
 */
#ifndef STSPIX_H
#define STSPIX_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  STSP_OK = 0,
  STSP_ERR_ARG = -1,      /* bad argument (null pointer, bad size) */
  STSP_ERR_CAPACITY = -2, /* output capacity smaller than the result */
  STSP_ERR_FILL = -3,     /* hole filling stopped shrinking (reference exits) */
  STSP_ERR_HIP = -4,      /* a HIP runtime call failed */
  STSP_ERR_STATE = -5,    /* stream step without a previous frame */
  STSP_ERR_TOO_LARGE = -6 /* frame exceeds stsp_bass_noalloc's static scratch buffer */
} stsp_status;

/* Algorithm settings; defaults are the reference's (pyapi.cu:63-73,
 * demo.cu:47-71).  Mirrors bist_forward's arguments plus the constants the
 * reference hard-codes. */
typedef struct {
  int niters;         /* outer iterations of frame 0 (bist_forward niters) */
  int niters_seg;     /* sweeps per iteration (pyapi.cu:63, 4) */
  int sm_start;       /* split/merge only when idx > sm_start (0) */
  int sp_size;        /* superpixel size (bist_forward sp_size) */
  int sm_period;      /* split at idx%p==0, merge at idx%p==p/2 (4) */
  int prop_niters;    /* >0: iterations of propagated frames; 0: adaptive */
  float sigma2_app;   /* bist_forward sigma2_app */
  float potts;        /* bist_forward potts */
  float alpha_hastings; /* bist_forward alpha (log alpha) */
  float split_alpha;  /* bist_forward split_alpha */
  float merge_alpha;  /* pyapi.cu:65, 0 */
  float thresh_relabel; /* pyapi.cu:68, 1e-5 */
  float thresh_new;   /* pyapi.cu:69, 1e-2 */
  int target_nspix;   /* >0: steer the split/merge alphas every second iteration
                         until the label count is within 5% of this (and at
                         least niters ran), at most 5000 iterations
                         (bass.cu:65-105, prop.cu:356-401; pyapi passes 0).
                         Single-frame calls only: the batch calls return
                         STSP_ERR_ARG for such contexts. */
} stsp_config;

/* Parameter arrays of one frame: the field set of the reference
 * SuperpixelParams (structs.h:13-48), device pointers, cap entries each
 * (x3 / x2 where noted). */
typedef struct {
  float* mu_app;              /* 3*cap */
  float* prior_mu_app;        /* 3*cap */
  double* mu_shape;           /* 2*cap */
  double* sigma_shape;        /* 3*cap, inverse covariance (xx, xy, yy) */
  float* logdet_sigma_shape;  /* cap */
  double* prior_mu_shape;     /* 2*cap */
  double* prior_sigma_shape;  /* 3*cap */
  double* sample_sigma_shape; /* 3*cap */
  int* counts;                /* cap */
  int* sm_counts;             /* cap */
  float* prior_counts;        /* cap */
  int* ids;                   /* cap */
  float* invalid_counts;      /* cap */
  int cap;
} stsp_params;

typedef struct stsp_ctx stsp_ctx;

/* Default configuration (reference values). */
void stsp_default_config(stsp_config* cfg);

/* A context owns the device workspace of one frame size; it replaces the
 * per-call cudaMalloc/cudaFree of run_bass / run_prop (bass.cu:233-264,
 * prop.cu:694-716). */
int stsp_ctx_create(const stsp_config* cfg, int height, int width, stsp_ctx** out);
void stsp_ctx_destroy(stsp_ctx* ctx);
int stsp_ctx_set_config(stsp_ctx* ctx, const stsp_config* cfg);

/* Launch shape of the fused boundary sweep (tile, sweeps per launch):
 * 0 = 32x16 x1, 1 = 32x32 x2, 2 = 64x32 x4, 3 = 64x64 x2, 4 = 64x32 x2,
 * 5 = 128x64 x2.  Results are identical for every shape; only speed
 * differs.  Returns STSP_ERR_ARG outside 0..5. */
int stsp_ctx_set_tuning(stsp_ctx* ctx, int sweep_shape);

/* Split-proposal search limits (no counterpart in the reference, whose
 * search is a level-synchronous flood fill over the whole frame,
 * split_merge_prop.cu:277-323).  Bounding boxes up to wave_box x wave_box
 * pixels (1..128) are searched by one wavefront; larger ones by a workgroup
 * in a window of at most win_rows x win_cols pixels (1..896, 1..512) around
 * the first seed, then over the whole box if the seed's connected component
 * leaves the window.  Results are identical for every setting; only speed
 * differs.  Defaults: 128, 896, 512. */
int stsp_ctx_set_split_search(stsp_ctx* ctx, int wave_box, int win_rows, int win_cols);

/* Shape (inverse-Wishart) term in the split Hastings ratio, off by default.
 * The reference computes it (marginal_likelihood_shape_p,
 * split_merge_prop.cu:1596-1621) but leaves its launches commented out
 * (split_merge_prop.cu:397-415), so the live reference ratio is the
 * appearance one alone; on = 1 adds, in FP64, for the two parts k, s of a
 * label against the whole f, L(k) + L(s) - L(f) with
 *   L = nu log det Psi - (nu + n) log det Sigma_hat
 * (Sigma_hat the sample covariance of the part's n pixel positions, Psi
 * the label's prior_sigma_shape, nu its prior count; L = -1e7 when
 * det Sigma_hat < 0.001; no term when det Psi < 0.001).  An extension:
 * results then differ from the reference's. */
int stsp_ctx_set_shape_marginal(stsp_ctx* ctx, int on);

/* Cumulative number of split-proposal searches by path since the context was
 * created: [0] one wavefront, [1] workgroup over the whole box, [2] workgroup
 * window that held the seed's component, [3] whole box after the component
 * left the window.  Waits for the device.  Diagnostic; no reference
 * counterpart. */
int stsp_ctx_split_search_counts(stsp_ctx* ctx, long long counts[4]);

/* Device bytes the context holds now (its workspace grows with the label
 * count of the frames it runs; a growth waits only for the context's own
 * previous call, not for the device).  For admission bookkeeping. */
int stsp_ctx_device_bytes(stsp_ctx* ctx, size_t* bytes);

/* Capacity an output parameter set needs: BASS 30*K0 (bass.cu:245),
 * propagated frame 10*nspix_prev (prop.cu:702). */
int stsp_bass_capacity(int sp_size, int height, int width);
int stsp_prop_capacity(int nspix_prev);

/* Single-image BASS: replaces run_bass (bass.cu:221-346).  Writes labels and
 * compacted parameters; *nspix_out = number of superpixels. */
int stsp_bass(stsp_ctx* ctx, const float* img, int* spix_out, stsp_params* out,
              int* nspix_out, hipStream_t stream);

/* One propagated frame: replaces shift_and_fill + run_prop
 * (shift_and_fill.cu:108-339, prop.cu:677-865) as called from main_loop
 * (pyapi.cu:128-170).  prev->prior_counts and prev->invalid_counts are
 * updated in place, as the reference does.  *niters_state receives the
 * adaptive iteration count (pyapi.cu:141-149). */
int stsp_prop(stsp_ctx* ctx, const float* img, const float* flow, const int* spix_prev,
              stsp_params* prev, int nspix_prev, int* spix_out, stsp_params* out,
              int* nspix_out, int* niters_state, hipStream_t stream);

/* stsp_prop plus the two intermediate maps the reference's run_prop and
 * stream_bass return as pre_fill / post_fill (st_spix/prop/__init__.py:
 * 282, 656): shifted_out receives the flow-shifted labels with -1 at holes
 * and collisions (run_shift_labels, shift_and_fill.cu:165-170), filled_out the
 * labels after hole filling and the connectivity pass, which start the
 * refinement (shift_and_fill.cu:188-209).  Either may be NULL; each holds
 * height*width int32 when given. */
int stsp_prop_ex(stsp_ctx* ctx, const float* img, const float* flow, const int* spix_prev,
                 stsp_params* prev, int nspix_prev, int* spix_out, stsp_params* out,
                 int* nspix_out, int* niters_state, int* shifted_out, int* filled_out,
                 hipStream_t stream);

/* Batched (served) form: the refinement loops of several propagated frames,
 * one per context, run as shared launches (one launch covers every frame;
 * per-label work is pulled from work lists), optionally replayed from a
 * captured hipGraph.  Each frame's result equals what stsp_prop returns for
 * it alone.  No counterpart in the reference, whose frame loop is sequential
 * (pyapi.cu:92-189).  Graph replay needs a non-default stream. */
typedef struct stsp_batch stsp_batch;
int stsp_batch_create(int max_frames, stsp_batch** out);
void stsp_batch_destroy(stsp_batch* batch);
int stsp_batch_use_graph(stsp_batch* batch, int on);
int stsp_prop_batch(stsp_batch* batch, stsp_ctx** ctxs, int n, const float** imgs,
                    const float** flows, const int** spix_prev, stsp_params* prevs,
                    const int* nspix_prev, int** spix_out, stsp_params* outs, int* nspix_out,
                    int* niters_state, hipStream_t stream);

/* Batched BASS over n images of any sizes (one context each): hexagonal
 * initialisation per image, then the refinement of all images in shared
 * launches (images with the same schedule together), then compaction and
 * export per image.  Per image the same result as stsp_bass.  Replaces a
 * loop of bass() calls over a set of images (bass.cu:178-330 per image;
 * indepent_bass, prop/__init__.py:38-85). */
int stsp_bass_batch(stsp_batch* batch, stsp_ctx** ctxs, int n, const float** imgs,
                    int** spix_out, stsp_params* outs, int* nspix_out, hipStream_t stream);

/* Single-image BASS that allocates nothing: the same computation and result
 * as stsp_bass (run_bass, bass.cu:221-346) for a context created with cfg at
 * height x width, with the whole workspace in a static device buffer of
 * 32 MiB (one copy per device, held until the process ends) instead of a
 * context.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Returns STSP_ERR_TOO_LARGE when the frame's workspace
 * (stsp_bass_noalloc_scratch_bytes) exceeds the buffer, STSP_ERR_ARG for
 * target_nspix > 0.  All calls on a device share the buffer: calls whose
 * device work overlaps race, so order them (one stream, or synchronise). */
int stsp_bass_noalloc(const stsp_config* cfg, int height, int width, const float* img,
                      int* spix_out, stsp_params* out, int* nspix_out, hipStream_t stream);
/* Scratch bytes stsp_bass_noalloc needs for this configuration and size. */
int stsp_bass_noalloc_scratch_bytes(const stsp_config* cfg, int height, int width,
                                    size_t* bytes);

/* Border overlay of get_marked_video (pyapi.cu:221-303,
 * find_border_pixels_end seg_utils.cu:218-257): marked = clamp(vid,0,1)
 * except label-border pixels, which take color[3]. nframes images. */
int stsp_marked_video(const float* vid, const int* spix, const float* color_host, int nframes,
                      int height, int width, float* marked, hipStream_t stream);

/* Colour preprocessing of the command-line path (demo.cu:248-250).
 * stsp_rescale_image: out = img_u8 / scale for 3*npix values per image
 * (rescale_image, rgb2lab.cu:19-32, on a device buffer).
 * stsp_rgb2lab: B,G,R floats in [0,1] -> Lab scaled to (L/-100, a/100,
 * b/100) (rgb2lab / rgb_to_lab, rgb2lab.cu:35-117).
 * stsp_lab2rgb: the inverse, writing R,G,B clamped to [0,1] (lab2rgb /
 * lab_to_rgb, rgb2lab.cu:42-162).  Pixels with a NaN channel are left
 * unwritten.  npix = pixels per image. */
int stsp_rescale_image(const uint8_t* img_u8, float* out, int nimages, int npix, float scale,
                       hipStream_t stream);
int stsp_rgb2lab(const float* img_bgr, float* img_lab, int nimages, int npix, hipStream_t stream);
int stsp_lab2rgb(const float* img_lab, float* img_rgb, int nimages, int npix, hipStream_t stream);

/* Device-time accounting of the hot kernels, measured with hipEvent pairs
 * on the launch stream around each launch of the class (no counterpart in
 * the reference, which prints host clock() times, demo.cu:243-340). */
enum { STSP_KCLASS_SWEEP = 0, STSP_KCLASS_STATS = 1, STSP_KCLASS_COUNT = 2 };
int stsp_ctx_enable_timing(stsp_ctx* ctx, int on);
int stsp_ctx_kernel_time(stsp_ctx* ctx, int kclass, double* total_ms, long long* launches);

/* Times the fused boundary sweep alone: after one (untimed) recount of the
 * label statistics and sweep records from the frames' final labels,
 * `nlaunch` launches back to back over the frames of the batch's last
 * refinement, between two hipEvents on `stream`; *ms_out receives the
 * elapsed time.  Each launch reads the final labels and writes the second
 * label buffer, so results already returned are unchanged; the working
 * statistics and parameters are rebuilt by the next frame's prologue.  For
 * benchmarks (no counterpart in the reference). */
int stsp_batch_time_sweeps(stsp_batch* batch, int nlaunch, hipStream_t stream, float* ms_out);

/* Sweep-shape autotuner probe (no counterpart in the reference, whose sweep
 * launch shape is fixed, update_seg.cu:281-311).  Over the frames of the
 * batch's last refinement, and like stsp_batch_time_sweeps leaving returned
 * results unchanged, times `nlaunch` launches of every fused-sweep launch
 * shape (tile width x height x sweeps per launch, STSP_SWEEP_SHAPES of
 * them, indices as stsp_ctx_set_tuning) after one untimed launch each;
 * us_per_sweep[STSP_SWEEP_SHAPES] receives microseconds per sweep and *best
 * the shape with the cheapest iteration (the sweeps per iteration of the
 * last refinement, as the refinement splits them into launches).  Blocks
 * until the probe is done.  All shapes give bit-identical labels. */
enum { STSP_SWEEP_SHAPES = 6 };
int stsp_batch_tune_sweeps(stsp_batch* batch, int nlaunch, hipStream_t stream,
                           float* us_per_sweep, int* best);

/* The per-step label ops of the st_spix.prop pipeline (the reference's
 * prop_cuda extension), standalone; the fused stsp_prop has its own shift
 * and fill (the BIST forms).  Device pointers; asynchronous on `stream`
 * unless stated.
 *
 * prop_cuda.shift_labels (csrc/prop/shift_labels.cu:81-126,
 * run_shift_labels(spix, flow) -> (shifted_spix, counts)): spix
 * [nbatch][H][W] int32; flow [nbatch][nspix][2] int32 as (dx, dy) per label.
 * Every pixel moves by its label's flow; each destination counts its
 * arrivals (counts) and keeps the largest arriving label (shifted; -1 where
 * none).  Destinations outside the image are dropped; pixels whose label is
 * outside [0, nspix) do not move (the reference reads past the flow array
 * there). */
int stsp_shift_labels(const int* spix, const int* flow, int nbatch, int height, int width,
                      int nspix, int* shifted, int* counts, hipStream_t stream);

/* prop_cuda.fill_missing (csrc/prop/fill_missing.cu:317-354,
 * run_fill_missing(spix, centers, missing, break_iter) -> filled): fills, in
 * place, the -1 pixels of seg [H][W] listed in missing[nmissing] (flat
 * indices; entries outside the image are ignored).  Each round marks the
 * listed -1 pixels with a labelled 4-neighbour and gives them, in four
 * checkerboard passes, the neighbouring label whose centre
 * centers[ncenters][2] (float x, y) is nearest in coordinates normalised by
 * width and height; labels >= ncenters are not candidates.  Rounds stop
 * when a round counts no -1 among the listed pixels, counts as many as the
 * round before (the first round compares with 1), or after break_iter
 * rounds (break_iter <= 0: no cap).  Blocks until done; *rounds_out
 * receives the rounds run and *negatives_out the listed -1 entries left
 * (the reference prints "negative spix exist." then).  Allocates a
 * workspace of H*W + 2.3 KB bytes on the stream. */
int stsp_fill_missing(int* seg, const float* centers, int ncenters, const int* missing,
                      int nmissing, int height, int width, int break_iter, int* rounds_out,
                      int* negatives_out, hipStream_t stream);

const char* stsp_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
