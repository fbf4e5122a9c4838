/* cuimg_hip — MI355X (gfx950) implementation of cuimg's frame-to-frame tracking
 * hot path: separable Gaussian, pyramid, FAST detector, pyramid Lucas-Kanade
 * particle tracker and its particle container, behind a C ABI.
 *
 * Conventions (same as the reference):
 *   - 2-vectors are (row, col); images are row-major with a pitch in BYTES and a
 *     border of `border` pixels around the domain; `begin` points at pixel (0,0)
 *     (cuimg/cpu/host_image2d.hpp:78-112, cuimg/gpu/kernel_image2d.h:17-85);
 *   - every function returns 0 on success, a negative cuimg error code, or a
 *     positive hipError_t; cuimg_hip_last_error() gives the message.  No C++
 *     exception crosses this boundary (the reference throws std::runtime_error
 *     from check_cuda_error, cuimg/error.h:17-28; the C++ wrapper rethrows);
 *   - device buffers passed in are plain device pointers; the library never
 *     takes ownership of caller memory;
 *   - a context owns one HIP stream; all work is enqueued on it in call order.
 *     A context (and the trackers made from it) is used by one host thread at a
 *     time, like the reference tracker.
 */
#ifndef CUIMG_HIP_H_
#define CUIMG_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUIMG_OK 0
#define CUIMG_E_INVALID (-1)     /* bad argument (shape, dtype, null pointer) */
#define CUIMG_E_NOMEM (-2)       /* device allocation failed */
#define CUIMG_E_UNSUPPORTED (-3) /* parameter combination not implemented */
#define CUIMG_E_CAPACITY (-4)    /* particle capacity exceeded */
#define CUIMG_E_NODEVICE (-5)    /* no HIP device */

enum cuimg_dtype {
  CUIMG_U8 = 1,    /* gl8u */
  CUIMG_F32 = 2,   /* gl01f / i_float1 */
  CUIMG_F32X4 = 3, /* i_float4 */
  CUIMG_U32 = 4,   /* unsigned int (sparse particle index) */
  CUIMG_F32X2 = 5, /* i_float2 (gradient) */
};

/* Non-owning image view (the kernel_image2d of the reference). */
typedef struct {
  void* begin;      /* device address of pixel (0,0) */
  int32_t nrows;
  int32_t ncols;
  int32_t pitch;    /* bytes between rows */
  int32_t border;   /* pixels of allocated border on every side */
  int32_t dtype;    /* enum cuimg_dtype */
} cuimg_img;

/* particle_f<translation_transform> (tracking2/particle_container.h:45-57):
 * 20 bytes, (row, col) vectors. */
typedef struct {
  float speed_r, speed_c;
  float pos_r, pos_c;
  uint16_t age;
  uint16_t pad;
} cuimg_particle_f;

typedef struct cuimg_hip_ctx cuimg_hip_ctx;

/* ---------------------------------------------------------------- context */
int cuimg_hip_ctx_create(int device, cuimg_hip_ctx** out);
/* Also destroys every tracker and batch made from the context that is still
 * alive (their handles are invalid afterwards); images allocated through the
 * context are freed. */
int cuimg_hip_ctx_destroy(cuimg_hip_ctx* ctx);
const char* cuimg_hip_last_error(const cuimg_hip_ctx* ctx);
int cuimg_hip_sync(cuimg_hip_ctx* ctx);
/* The context's hipStream_t. */
void* cuimg_hip_stream(cuimg_hip_ctx* ctx);
/* Per-kernel timing with HIP events recorded on the context stream. */
int cuimg_hip_profile_enable(cuimg_hip_ctx* ctx, int on);
int cuimg_hip_profile_reset(cuimg_hip_ctx* ctx);
/* Aggregated (launch count, total milliseconds) of the kernel named `kernel`;
 * synchronises the context. */
int cuimg_hip_profile_query(cuimg_hip_ctx* ctx, const char* kernel, int64_t* launches,
                            double* total_ms);
/* Span timer: milliseconds between the two events recorded on the context
 * stream by start and stop (stop synchronises); no counterpart in the
 * reference, used to time back-to-back launches without per-launch events. */
int cuimg_hip_timer_start(cuimg_hip_ctx* ctx);
int cuimg_hip_timer_stop(cuimg_hip_ctx* ctx, double* ms);

/* --------------------------------------------------------- memory / images */
int cuimg_hip_malloc(cuimg_hip_ctx* ctx, size_t bytes, void** out);
int cuimg_hip_free(cuimg_hip_ctx* ctx, void* p);
int cuimg_hip_memcpy_h2d(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int cuimg_hip_memcpy_d2h(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy, asynchronous on the context stream (the copy-kernel
 * reference for bandwidth and traffic measurements). */
int cuimg_hip_memcpy_d2d(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes);
int cuimg_hip_memset(cuimg_hip_ctx* ctx, void* dst, int value, size_t bytes);

/* Replaces device_image2d<V>(domain, border) (cuimg/gpu/device_image2d.hpp:23-38):
 * rows are 128-byte aligned at pixel (0,0) and the pitch is a multiple of 128. */
int cuimg_hip_image_alloc(cuimg_hip_ctx* ctx, int nrows, int ncols, int border, int dtype,
                          cuimg_img* out);
int cuimg_hip_image_free(cuimg_hip_ctx* ctx, cuimg_img* img);
/* copy(host, device) / copy(device, host) of the domain (cuimg/copy.h:31-71). */
int cuimg_hip_image_upload(cuimg_hip_ctx* ctx, const cuimg_img* dst, const void* host,
                           size_t host_pitch);
int cuimg_hip_image_download(cuimg_hip_ctx* ctx, const cuimg_img* src, void* host,
                             size_t host_pitch);
/* fill_border_clamp (cuimg/border.h:128-137). */
int cuimg_hip_fill_border_clamp(cuimg_hip_ctx* ctx, const cuimg_img* img);

/* --------------------------------------------------------------- Gaussian */
/* meta_kernel<meta_gaussian<0, sigma>, x>::coef() for x = -half_width..half_width
 * (generator tools/make_meta_gaussian/make_meta_gaussian.cpp:63-105). Host only. */
int cuimg_hip_gaussian_coefs(int sigma, int half_width, float* out);
/* local_jet_static<.., 0, 0, sigma, half_width>(in, out, tmp)
 * (cuimg/gpu/local_jet_static.h:13-28) = meta_convolve_row2d then col2d
 * (cuimg/gpu/meta_convolve.h:131-161, 288-314): clamp-to-edge, taps
 * x = -K..K.  One fused pass, no temporary image.  in: U8, F32 or F32X4;
 * out: F32 (F32X4 for F32X4 input); same domain.  half_width <= 12. */
int cuimg_hip_gaussian_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           int sigma, int half_width);
/* Same with explicit coefficients (2*half_width+1 floats). */
int cuimg_hip_convolve_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           const float* coefs, int half_width);
/* Several independent blurs in one launch (e.g. all pyramid levels). */
int cuimg_hip_gaussian_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs,
                                 int n, int sigma, int half_width);

/* ---------------------------------------------------------------- pyramid */
enum {
  /* subsample (cuimg/gpu/mipmap.h:206-227, kernel :25-50): 2x2 box mean, integer
   * division for U8.  U8 or F32 in and out. */
  CUIMG_PYR_BOX2X2 = 0,
  /* build_mipmap's blur-then-subsample level (cuimg/gpu/mipmap.h:94-109,
   * disabled in the reference): Gaussian sigma 1, K = 6 with clamp-to-edge,
   * sampled at 2p.  U8 or F32 in, F32 out. */
  CUIMG_PYR_GAUSS_DECIMATE = 1
};
/* One pyramid level: out domain = in / 2 (floor); mode is a CUIMG_PYR_* value. */
int cuimg_hip_pyr_down(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out, int mode);
/* tracker::update_input (tracking2/tracker.hpp:53-69) fused: copies `frame`
 * into levels[0], builds levels[1..n-1] with pyr_down and fills every level's
 * border by clamping.  U8 (frame may be levels[0] itself) or F32 (no border
 * fill needed: pass border 0). */
int cuimg_hip_pyramid_build(cuimg_hip_ctx* ctx, const cuimg_img* frame, const cuimg_img* levels,
                            int nlevels);

/* ------------------------------------------------------------------- FAST */
/* compute_saliency_kernel + fast_detector::update
 * (tracking2/fast_detector.hpp:26-157, 267-313): out <- 0, then the FAST score
 * (modulo 256) wherever mask != 0 (mask may be NULL = everywhere).  in: U8 with
 * border >= 3 filled by clamping; mask, out: U8 of the same domain. */
int cuimg_hip_fast_saliency(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* mask,
                            const cuimg_img* out, int n, int th);

/* Detector mask (create_detector_mask + particle_mask_kernel,
 * tracking2/pyrlk_tracker.hpp:879-888, tracking2/tracking_strategies.hpp:201-225):
 * mask <- 255, then 0 on the 3x3 block around the truncated position of each of
 * the n particles (device array, dead particles included); pixels outside the
 * mask's domain are left alone.  Ordered on the context's stream. */
int cuimg_hip_particle_mask(cuimg_hip_ctx* ctx, const cuimg_particle_f* parts, uint32_t n,
                            const cuimg_img* mask);

/* Box selection of new points (fast_detector::new_particles,
 * tracking2/fast_detector.hpp:217-262): one box of box x box pixels at
 * (i*box + 1, j*box + 1) per cell of domain/box; a box holding a particle
 * (sparse != 0xFFFFFFFF) is skipped; its first strict saliency maximum in raster
 * order is kept unless below th; the kept points inside domain - border(3) are
 * written as (row, col) int32 pairs to the device array pts in raster order
 * (the order particle_container::add receives them).  *count = number of
 * points (only the first cap are written).  sal: U8, sparse: U32, same domain;
 * box in 1..15.  Synchronises the context. */
int cuimg_hip_fast_select(cuimg_hip_ctx* ctx, const cuimg_img* sal, const cuimg_img* sparse, int box,
                          float th, int32_t* pts, uint32_t cap, uint32_t* count);

/* ---------------------------------------------------------------- tracker */
typedef struct {
  int32_t nscales;
  int32_t profile;              /* 0: reference matcher on u8 frames (ws 11, 30 it,
                                   top-25); 1: blurred FP32 frames; 2: the bc2s
                                   descriptor with the gradient-descent matcher
                                   (generic_strategy, tracking_strategies.hpp:28-111;
                                   the LK and blur settings are unused) */
  int32_t lk_ws;                /* window side (odd) */
  int32_t lk_max_iters;
  float lk_eps;                 /* stop when |delta| < eps */
  int32_t lk_top_n;             /* strongest-gradient pixels; >= ws*ws = all */
  int32_t fast_n;               /* fast_detector::set_n */
  float fast_th;                /* fast_detector::set_fast_threshold */
  int32_t box_size;             /* fast_detector::set_box_size */
  int32_t detector_frequency;   /* pyrlk_cpu::set_detector_frequency */
  int32_t filtering_frequency;  /* pyrlk_cpu::set_filtering_frequency */
  int32_t k;                    /* pyrlk_cpu::set_k */
  int32_t with_merge;           /* pyrlk_cpu::set_with_merge */
  int32_t blur_sigma;           /* profile 1 blur */
  int32_t blur_K;
} cuimg_tracker_params;

typedef struct cuimg_hip_tracker cuimg_hip_tracker;

/* Settings the reference keeps per scale (one pyrlk_cpu strategy and one
 * fast_detector per scale): fast_detector::set_n / set_fast_threshold /
 * set_box_size (tracking2/fast_detector.hpp:170-199), pyrlk_cpu::
 * set_detector_frequency / set_filtering_frequency / set_k / set_with_merge /
 * set_particle_domain (pyrlk_tracker.h:45-58, pyrlk_tracker.hpp:894-897). */
typedef struct {
  int32_t fast_n;
  float fast_th;
  int32_t detector_frequency;
  int32_t filtering_frequency;
  int32_t k;
  int32_t with_merge;
  int32_t box_size;             /* 1..15 */
  /* particle domain, an inclusive (row, col) box: particles outside it are
   * removed and matches must land in it; default domain - border(3) */
  int32_t pd_r1, pd_c1, pd_r2, pd_c2;
} cuimg_scale_params;

/* tracker<pyrlk_cpu<fast_detector<cpu>, iterative_lk_match_5x5>>(domain, nscales)
 * (tracking2/tracker.hpp:12-26, pyrlk_tracker.hpp:771-793). */
int cuimg_hip_tracker_create(cuimg_hip_ctx* ctx, int nrows, int ncols,
                             const cuimg_tracker_params* params, cuimg_hip_tracker** out);
int cuimg_hip_tracker_destroy(cuimg_hip_tracker* t);
/* tracker::run(frame) (tracking2/tracker.hpp:73-80): one frame step over all
 * scales.  `frame` is a U8 device image of the tracker domain (any border).
 * Asynchronous on the context stream. */
int cuimg_hip_tracker_run(cuimg_hip_tracker* t, const cuimg_img* frame);
/* Same from a host gray frame (pitch in bytes).  The frame is copied into one
 * of two pinned staging buffers before the call returns (the caller may reuse
 * it at once) and uploaded on a copy stream while the previous step still
 * runs; the step waits for its upload. */
int cuimg_hip_tracker_run_host(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch);
/* Same from a host frame of `channels` interleaved u8 channels (1, 3 or 4):
 * the gray frame is channel 0, as the reference sample's get_x(frame)
 * (samples/tracking/main.cc:116-117), extracted on the device after the
 * upload of the packed frame. */
int cuimg_hip_tracker_run_host_rgb(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch, int channels);
/* Per-scale settings; take effect from the next frame.  Reading fills `out`. */
int cuimg_hip_tracker_set_scale_params(cuimg_hip_tracker* t, int scale, const cuimg_scale_params* in);
int cuimg_hip_tracker_get_scale_params(cuimg_hip_tracker* t, int scale, cuimg_scale_params* out);
/* tracker::clear (tracking2/tracker.hpp:95-102). */
int cuimg_hip_tracker_clear(cuimg_hip_tracker* t);
/* pyrlk_cpu::clear of scale s (pyrlk_tracker.hpp:1015-1019): the scale's
 * frame counter restarts at 0 (its next frame matches nothing and detects);
 * its particles stay. */
int cuimg_hip_tracker_clear_scale(cuimg_hip_tracker* t, int scale);
/* tracker::update_input (tracking2/tracker.hpp:53-69): the pyramid of the
 * next frame (device or host gray frame, staged like tracker_run_host), and
 * tracker::run() (:83-91): the frame step of every scale on the input last
 * staged (on the previous input again when none was staged since the last
 * step).  tracker_run = update_input + step, in one graph replay. */
int cuimg_hip_tracker_update_input(cuimg_hip_tracker* t, const cuimg_img* frame);
int cuimg_hip_tracker_update_input_host(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch);
int cuimg_hip_tracker_step(cuimg_hip_tracker* t);
/* Number of entries of scale(s).pset().dense_particles(); synchronises. */
int cuimg_hip_tracker_nparticles(cuimg_hip_tracker* t, int scale, uint32_t* n);
/* Copies dense_particles() of scale s to host (cap entries); synchronises. */
int cuimg_hip_tracker_get_particles(cuimg_hip_tracker* t, int scale, cuimg_particle_f* out,
                                    uint32_t cap, uint32_t* n);
/* Replaces dense_particles() of scale s (for replay and tests). */
int cuimg_hip_tracker_set_particles(cuimg_hip_tracker* t, int scale, const cuimg_particle_f* in,
                                    uint32_t n);
/* matches() (old index -> new index, 0xFFFFFFFF = died) of the last compaction. */
int cuimg_hip_tracker_get_matches(cuimg_hip_tracker* t, int scale, uint32_t* out, uint32_t cap,
                                  uint32_t* n);
int cuimg_hip_tracker_compact_has_run(cuimg_hip_tracker* t, int scale, int* out);
/* sparse_particles(): U32 image of the scale domain, 0xFFFFFFFF = empty. */
int cuimg_hip_tracker_sparse(cuimg_hip_tracker* t, int scale, cuimg_img* view);
/* input() of scale s: U8 level with border 3. */
int cuimg_hip_tracker_level(cuimg_hip_tracker* t, int scale, cuimg_img* view);
/* detector().saliency() of scale s. */
int cuimg_hip_tracker_saliency(cuimg_hip_tracker* t, int scale, cuimg_img* view);
/* Per-particle matcher decisions of the last frame at scale s (before the
 * container update): accept flag, translation (row, col), distance. */
int cuimg_hip_tracker_get_decisions(cuimg_hip_tracker* t, int scale, int32_t* accept, float* tr2,
                                    float* dist, uint32_t cap, uint32_t* n);
/* Matcher images of scale s: previous / current frame as F32 and the F32X2
 * gradient of the previous frame. */
int cuimg_hip_tracker_matcher_images(cuimg_hip_tracker* t, int scale, cuimg_img* A, cuimg_img* B,
                                     cuimg_img* G);
int cuimg_hip_tracker_frame_count(cuimg_hip_tracker* t, int* out);
/* Totals since creation (or clear) at scale s: particles with age > 0 run
 * through the matcher, and how many of them were accepted.  Synchronises. */
int cuimg_hip_tracker_work(cuimg_hip_tracker* t, int scale, uint64_t* live, uint64_t* accepted);

/* Profile 2 (bc2s strategy).  Particle positions and speeds are integers (the
 * reference's particle::pos is i_short2) held in the float fields.
 * The 16-byte descriptors of the dense vector of scale s
 * (particle_container::features(), tracking2/particle_container.h:95-97):
 * bytes 0..7 from s1, 8..15 from s2 (the reference's weights are 255 for
 * every descriptor and are not stored).  Synchronises. */
int cuimg_hip_tracker_get_features(cuimg_hip_tracker* t, int scale, uint8_t* out, uint32_t cap, uint32_t* n);
/* Views of the scale's feature images s1 / s2 (bc2s_feature::s1(), s2(),
 * tracking2/bc2s_feature.h:76-77; borders 3 / 6, clamp-filled) and of its
 * contrast (contrast_ of generic_strategy, saturated at 255). */
int cuimg_hip_tracker_bc2s_images(cuimg_hip_tracker* t, int scale, cuimg_img* s1, cuimg_img* s2,
                                  cuimg_img* contrast);
/* Views of the scale's flow (generic_strategy::flow(), tracking_strategies.h:82;
 * F32X2 per 8x8 cell, (-500, -500) = no flow) and multiscale count (U32 per
 * cell) of the last frame. */
int cuimg_hip_tracker_flow(cuimg_hip_tracker* t, int scale, cuimg_img* flow, cuimg_img* msc);

/* bc2s_feature<cpu>::update (tracking2/bc2s_feature.hpp:232-253, the build
 * without OpenCV): s1 = blur121(in), s2 = blur121(s1), each border of s1 / s2
 * (up to 3 / 6) clamp-filled; and contrast_kernel
 * (tracking2/tracking_strategies.hpp:104-136) into `contrast` (U32, or U8
 * saturated at 255, or NULL).  in: U8 of any border (reads clamp-to-edge). */
int cuimg_hip_bc2s_images(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* s1, const cuimg_img* s2,
                          const cuimg_img* contrast);
/* gaussian_blur_sigma1 (cpu/gaussian_blur.h:9-47): (2 v + left + right) / 4 on
 * the rows, then on the columns, in integers; the edge pixels the reference
 * leaves unwritten use the edge pixel for the missing neighbour. */
int cuimg_hip_blur121(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out);
/* gaussian_blur(src, dst, tmp, gaussian_kernel<cuda_gpu>(sigma, hw))
 * (gpu/gaussian_blur.h:31-93): taps (int)(100 g(sigma, i)), i = -hw..hw; a
 * column pass into a u8 temporary, then a row pass, each divided by the tap
 * sum with truncation, centred, clamp-to-edge.  hw <= 15.  Synchronises. */
int cuimg_hip_gaussian_int(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out, float sigma, int hw);
/* bc2s_feature::operator() (tracking2/bc2s_feature.hpp:119-143) at n points
 * (device (row, col) int32 pairs) into out16 (device, 16 bytes each).  s1 / s2
 * as the tracker holds them (borders >= 3 / >= 6, clamp-filled). */
int cuimg_hip_bc2s_features(cuimg_hip_ctx* ctx, const cuimg_img* s1, const cuimg_img* s2, const int32_t* pts,
                            uint32_t n, uint8_t* out16);
/* two_step_gradient_descent_match (scale 0) or gradient_descent_match at
 * scale 1 / 2 (tracking2/gradient_descent_matcher.hpp:9-125) of n descriptors
 * (device, 16 bytes each) from n predictions (device (row, col) int32 pairs):
 * displacements (device int32 pairs) and distances (device floats). */
int cuimg_hip_bc2s_match(cuimg_hip_ctx* ctx, const cuimg_img* s1, const cuimg_img* s2, const uint8_t* feats16,
                         const int32_t* preds, uint32_t n, int scale, int32_t* out_disp, float* out_dist);

/* merge_trajectories_kernel over a dense vector (tracking2/merge_trajectories.h:10-35,
 * launched at pyrlk_tracker.hpp:865-870), with the result of the serial
 * ascending-index loop: `parts` (device, n entries) is updated in place,
 * `sparse` is the U32 index image of the domain (0xFFFFFFFF = empty).
 * Synchronises the context. */
int cuimg_hip_particles_merge(cuimg_hip_ctx* ctx, cuimg_particle_f* parts, uint32_t n,
                              const cuimg_img* sparse);

/* particle_container::add (tracking2/particle_container.hpp:370-387) for each
 * point of pts (device array of npts (row, col) int32 pairs) in order, while the
 * count stays below cap: particle n + k gets age 1, zero speed and the point as
 * position, and sparse(point) = n + k (the last of equal points wins; points
 * outside the sparse domain set no index).  parts: device array of cap entries;
 * *n_out = n + number added.  Ordered on the context's stream. */
int cuimg_hip_particles_add(cuimg_hip_ctx* ctx, cuimg_particle_f* parts, uint32_t n, uint32_t cap,
                            const int32_t* pts, uint32_t npts, const cuimg_img* sparse, uint32_t* n_out);

/* particle_container::compact(cpu) (tracking2/particle_container.hpp:216-252):
 * the particles of `parts` with age != 0, in order, are written to `out` (a
 * separate device array of n entries); matches[i] = new index of particle i or
 * 0xFFFFFFFF; every kept particle's truncated position gets its new index in the
 * U32 `sparse` image (the last kept particle of a pixel wins; other pixels are
 * unchanged).  *n_out = kept count.  Synchronises the context. */
int cuimg_hip_particles_compact(cuimg_hip_ctx* ctx, const cuimg_particle_f* parts, uint32_t n,
                                cuimg_particle_f* out, uint32_t* matches, const cuimg_img* sparse,
                                uint32_t* n_out);

/* Frame steps are captured once into a hipGraph and replayed (default on);
 * disabled, every step launches its kernels (per-kernel timing needs this). */
int cuimg_hip_tracker_set_graph(cuimg_hip_tracker* t, int enable);
/* Launch shape of a tracker's frame step (see cuimg_hip_batch_set_tuning). */
int cuimg_hip_tracker_set_tuning(cuimg_hip_tracker* t, int lk_ppw, int blur_th);

/* --------------------------------------------------------------- matcher */
/* iterative_lk_match_5x5 settings (tracking2/pyrlk_tracker.hpp:114-128). */
typedef struct {
  int32_t ws;           /* window side (odd, <= 15) */
  int32_t max_iters;
  float eps;            /* stop when |delta| < eps */
  int32_t top_n;        /* strongest-gradient pixels; >= ws*ws = all */
  int32_t k;            /* accept when err < k (pyrlk_cpu::set_k) */
  int32_t template_u8;  /* template samples stored as uint8 (the reference's u8 frames) */
} cuimg_lk_params;

/* pyrlk_match_particles_kernel (tracking2/pyrlk_tracker.hpp:676-769) for the n
 * particles of scale `scale`, with the recursive coarse-to-fine prediction over
 * levels scale+1 .. nlevels-1 (predictions.h:60-74): prev/cur are the level
 * images (U8 or F32, nlevels each) of the previous and current frame.  An
 * accepted particle is moved (age+1, speed, position) and, when `sparse` is
 * given (U32 image of the scale's domain), written there with the last index
 * winning; a rejected one gets age 0.  accept/tr2/dist (device, n entries, may
 * all be NULL) receive the decisions.  parts are device memory.  Synchronises. */
int cuimg_hip_lk_track(cuimg_hip_ctx* ctx, const cuimg_img* prev, const cuimg_img* cur, int nlevels,
                       int scale, cuimg_particle_f* parts, uint32_t n, const cuimg_lk_params* params,
                       const cuimg_img* sparse, int32_t* accept, float* tr2, float* dist);
/* cuimg_hip_lk_track without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results.  Its scratch (the matcher's counters and, when
 * accept / tr2 / dist are NULL, the decision buffers) is a fixed device array
 * of 256 B + 16 B x 262144 particles, one per GPU, held until the process
 * ends: with NULL decision buffers n is at most 262144 (CUIMG_E_CAPACITY
 * beyond).  Every call on a GPU uses that GPU's array: calls that overlap on
 * one GPU (other host threads or contexts) race. */
int cuimg_hip_lk_track_noalloc(cuimg_hip_ctx* ctx, const cuimg_img* prev, const cuimg_img* cur, int nlevels,
                               int scale, cuimg_particle_f* parts, uint32_t n, const cuimg_lk_params* params,
                               const cuimg_img* sparse, int32_t* accept, float* tr2, float* dist);

/* ----------------------------------------------------------------- batch */
/* Many independent trackers (video streams, each its own domain and nscales;
 * one matcher profile / window / blur for all) stepped together: one frame
 * step of every stream is one fixed sequence of batched launches, captured
 * once into a hipGraph.  No counterpart in the reference, which runs one
 * tracker<F> per stream (tracking2/tracker.hpp:73-91); each stream's results
 * equal those of its own tracker. */
typedef struct cuimg_hip_batch cuimg_hip_batch;
int cuimg_hip_batch_create(cuimg_hip_ctx* ctx, int nstreams, const int* nrows, const int* ncols,
                           const cuimg_tracker_params* params, cuimg_hip_batch** out);
int cuimg_hip_batch_destroy(cuimg_hip_batch* b);
/* One frame step of every stream: frames[i] is a U8 device image of stream i's
 * domain.  Asynchronous on the context stream. */
int cuimg_hip_batch_step(cuimg_hip_batch* b, const cuimg_img* frames);
/* The same from host frames: frames[i] (pitches[i] bytes per row) holds stream
 * i's frame with `channels` interleaved u8 channels (1, 3 or 4; channel 0 is
 * the gray frame).  Each is staged like cuimg_hip_tracker_run_host: copied into
 * a pinned buffer before the call returns, uploaded while the previous step
 * runs.  No counterpart in the reference (which has no multi-stream form). */
int cuimg_hip_batch_step_host(cuimg_hip_batch* b, const uint8_t* const* frames, const size_t* pitches,
                              int channels);
int cuimg_hip_batch_set_graph(cuimg_hip_batch* b, int enable);
/* Tracker view of stream i for the tracker accessors (owned by the batch;
 * tracker_run on it is refused, tracker_destroy does nothing). */
int cuimg_hip_batch_tracker(cuimg_hip_batch* b, int stream, cuimg_hip_tracker** out);
/* Launch shape: lk_ppw = particles per matcher wave (a workgroup of 4 waves
 * takes 4 * lk_ppw consecutive particles; default 4), blur_th = blur tile
 * height in rows (0 = chosen from the total work).  Reading reports the
 * values in use. */
int cuimg_hip_batch_set_tuning(cuimg_hip_batch* b, int lk_ppw, int blur_th);
int cuimg_hip_batch_get_tuning(cuimg_hip_batch* b, int* lk_ppw, int* blur_th);
int cuimg_hip_batch_clear(cuimg_hip_batch* b);

/* ------------------------------------------------------------- adjoint */
/* Photometric LK loss of the blurred frames and its gradient (no counterpart in
 * the reference; it is the term iterative_lk_match_5x5 minimises,
 * tracking2/pyrlk_tracker.hpp:283-317, written as a loss):
 *   L = sum_p sum_w ((G*I1)(x_p + d_p + w) - (G*I0)(x_p + w))^2
 * G = gaussian_sep(sigma, half_width), bilinear samples as the matcher's li
 * (:130-146), w over a ws x ws window.  Device inputs: pos2, d2 (n float2,
 * (row, col)).  Device outputs: *loss_total (FP64), loss_p[n], dL_dd2[n float2] and,
 * if dL_dI1 is not NULL, dL/dI1 (F32 image of the frame domain).  Blurred
 * values are recomputed per particle, never stored.  scratch: device buffer of
 * at least 2 * nrows * ncols floats.  ws <= 7 (odd), half_width <= 6. */
int cuimg_hip_lk_loss_adjoint(cuimg_hip_ctx* ctx, const cuimg_img* I0, const cuimg_img* I1,
                              const float* pos2, const float* d2, uint32_t n, int sigma,
                              int half_width, int ws, double* loss_total, float* loss_p,
                              float* dL_dd2, const cuimg_img* dL_dI1, float* scratch,
                              size_t scratch_floats);

/* -------------------------------------------------------- synthetic input */
/* Seeded texture video used by the benchmarks (host code, OpenMP).  motion:
 * 0 = translation U[-3,3]^2 per frame, 1 = affine (rotation +-2 deg, scale
 * [0.98,1.02], translation +-8 px).  out: nframes * nrows * ncols bytes. */
int cuimg_synth_video(uint64_t seed, int nrows, int ncols, int nframes, int motion, uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
