/* C entry points of the CPU oracle (loaded with ctypes by tests/ and bench.py's
 * cpu_baseline leg).  TEST INFRASTRUCTURE ONLY: the product never calls these.
 * Images are dense row-major arrays without border; 2-vectors are (row, col). */
#ifndef CUIMG_ORACLE_H_
#define CUIMG_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float speed_r, speed_c;
  float pos_r, pos_c;
  uint16_t age;
  uint16_t pad;
} orc_particle;

typedef struct {
  int32_t nscales;
  int32_t profile;        /* 0: reference defaults (u8), 1: blurred FP32, 2: bc2s */
  int32_t lk_ws;
  int32_t lk_max_iters;
  float lk_eps;
  int32_t lk_top_n;
  int32_t fast_n;
  float fast_th;
  int32_t box_size;
  int32_t detector_frequency;
  int32_t filtering_frequency;
  int32_t k;
  int32_t with_merge;
  int32_t blur_sigma;
  int32_t blur_K;
  int32_t sum_lanes;      /* order of the matcher's window sums: 0 raster, 16 / 64 lane groups */
} orc_tracker_params;

float orc_gauss_coef_runtime(float sigma, int x);
float orc_gauss_coef_meta(int sigma, int x);
void orc_gauss_coefs(int sigma, int K, int meta_rounding, float* out);

/* Separable Gaussian, GPU meta-kernel semantics; ch interleaved channels. */
/* fused: taps added with fused multiply-adds (the GPU's arithmetic). */
void orc_gaussian_sep_f32(const float* in, int nr, int nc, int ch, const float* coef, int K, int fused,
                          float* out);
void orc_gaussian_sep_u8(const uint8_t* in, int nr, int nc, const float* coef, int K, int fused, float* out);
void orc_gaussian_sep_f64(const double* in, int nr, int nc, int ch, const double* coef, int K,
                          double* out);

void orc_pyr_down_u8(const uint8_t* in, int nr, int nc, uint8_t* out);
void orc_pyr_down_f32(const float* in, int nr, int nc, float* out);

uint8_t orc_fast_score(const uint8_t* in, int nr, int nc, int r, int c, int n, int th);
void orc_fast_saliency(const uint8_t* in, int nr, int nc, const uint8_t* mask, int n, int th,
                       uint8_t* out);

void orc_sobel_gradient(const float* prev, int nr, int nc, int integer_source, float* out2);

/* One LK match on a single level (A, B float images, G = 2-channel gradient). */
void orc_lk_match(const float* A, const float* B, const float* G, int nr, int nc, int ws,
                  int max_iters, float eps, int top_n, int template_u8, int sum_lanes, float p_r,
                  float p_c, float pred_r, float pred_c, float* out_tr2, float* out_dist);

void* orc_tracker_create(int nr, int nc, const orc_tracker_params* p);
void orc_tracker_destroy(void* t);
void orc_tracker_clear(void* t);
void orc_tracker_run(void* t, const uint8_t* frame);
/* tracker::update_input and tracker::run() (tracking2/tracker.hpp:53-91). */
void orc_tracker_update_input(void* t, const uint8_t* frame);
void orc_tracker_step(void* t);
/* pyrlk_cpu::clear of one scale (frame counter only, pyrlk_tracker.hpp:1015-1019). */
void orc_tracker_clear_scale(void* t, int s);
/* fast_detector::set_box_size and pyrlk_cpu::set_particle_domain of one scale. */
void orc_tracker_set_scale(void* t, int s, int box_size, int r1, int c1, int r2, int c2);
int orc_tracker_nscales(void* t);
void orc_tracker_level_dims(void* t, int s, int* nr, int* nc);
int orc_tracker_nparticles(void* t, int s);
int orc_tracker_get_particles(void* t, int s, orc_particle* out, int cap);
int orc_tracker_set_particles(void* t, int s, const orc_particle* in, int n);
int orc_tracker_get_matches(void* t, int s, uint32_t* out, int cap);
int orc_tracker_compact_has_run(void* t, int s);
void orc_tracker_get_sparse(void* t, int s, int32_t* out);
void orc_tracker_get_saliency(void* t, int s, uint8_t* out);
void orc_tracker_get_level(void* t, int s, uint8_t* out);
void orc_tracker_get_matcher_images(void* t, int s, float* A, float* B, float* G2);
/* Last frame's per-particle LK decision (before move/remove) at scale s. */
int orc_tracker_get_decisions(void* t, int s, int32_t* accept, float* tr2, float* dist, int cap);

/* Per-particle decisions of one scale's matcher on the current state, without
 * changing it: lets a test feed identical particles to the GPU. */
int orc_tracker_lk_decide(void* t, int s, int32_t* accept, float* tr2, float* dist, int cap);

/* merge_trajectories over a container state (in place); sparse: nr*nc, -1 empty. */
void orc_merge(orc_particle* parts, int n, const int32_t* sparse, int nr, int nc);
/* Detector mask of a dense vector: 255, 0 on the 3x3 block of every particle
 * (tracking_strategies.hpp:201-225). */
void orc_particle_mask(const orc_particle* parts, int n, int nr, int nc, uint8_t* mask);
/* Box selection of new points (fast_detector.hpp:217-262), (row, col) pairs in
 * raster order; returns the number of points (at most cap written). */
int orc_fast_select(const uint8_t* sal, const int32_t* sparse, int nr, int nc, int box, float th, int32_t* pts,
                    int cap);

/* ------------------------------------------------------------ bc2s path */
/* Integer 1-2-1 blur, rows then columns, each / 4 truncated, clamp-to-edge
 * (cpu/gaussian_blur.h:9-47). */
void orc_blur121_u8(const uint8_t* in, int nr, int nc, uint8_t* out);
/* Taps (int)(100 g(sigma, i)), i = -hw..hw (gpu/gaussian_blur.h:31-53). */
void orc_gaussian_int_taps(float sigma, int hw, int32_t* out);
/* Separable integer Gaussian with those taps (columns then rows, each pass
 * divided by the tap sum with truncation into u8, clamp-to-edge). */
void orc_gaussian_int_u8(const uint8_t* in, int nr, int nc, float sigma, int hw, uint8_t* out);
/* contrast_kernel (tracking_strategies.hpp:104-136), clamp-to-edge reads. */
void orc_contrast_u32(const uint8_t* in, int nr, int nc, uint32_t* out);
/* bc2s descriptor at (r, c) of the feature images s1, s2 (16 bytes). */
void orc_bc2s_feature(const uint8_t* s1, const uint8_t* s2, int nr, int nc, int r, int c, uint8_t* out16);
int orc_bc2s_distance(const uint8_t* s1, const uint8_t* s2, int nr, int nc, const uint8_t* f16, int r, int c,
                      int scale);
/* scale 1 or 2: gradient_descent_match at that scale; 0: the two-step match.
 * out_d2: (row, col) displacement from the prediction. */
void orc_bc2s_match(const uint8_t* s1, const uint8_t* s2, int nr, int nc, const uint8_t* f16, int pred_r,
                    int pred_c, int scale, int32_t* out_d2, float* out_dist);
/* Tracker state of the bc2s strategy (profile 2). */
int orc_tracker_get_features(void* t, int s, uint8_t* out16, int cap);
void orc_tracker_get_bc2s_images(void* t, int s, uint8_t* s1, uint8_t* s2, uint32_t* contrast);
void orc_tracker_flow_dims(void* t, int s, int* cr, int* cc);
void orc_tracker_get_flow(void* t, int s, float* flow2, uint32_t* msc);

/* OpenMP threads of the oracle's parallel loops. */
void orc_set_threads(int n);
int orc_max_threads(void);

/* particle_container::add for up to cap - n points (pts: npts (row, col)
 * pairs), in place; returns the new particle count (particle_container.hpp:370-387). */
int orc_particles_add(orc_particle* parts, int n, int cap, int32_t* sparse, int nr, int nc, const int32_t* pts,
                      int npts);
/* particle_container::compact(cpu) in place (particle_container.hpp:216-252);
 * matches: n entries; returns the kept count. */
int orc_particles_compact(orc_particle* parts, int n, int32_t* sparse, int nr, int nc, uint32_t* matches);

/* FP64 photometric LK loss and its gradient (blurred frames, bilinear samples). */
double orc_lk_loss_grad_f64(const float* I0, const float* I1, int nr, int nc, const float* coef,
                            int K, const float* pos2, const float* d2, int np, int ws,
                            double* dL_dd2, double* dL_dI1);
/* Same in FP64 inputs (used for central finite differences). */
double orc_lk_loss_grad_f64_d(const double* I0, const double* I1, int nr, int nc,
                              const double* coef, int K, const double* pos2, const double* d2,
                              int np, int ws, double* dL_dd2, double* dL_dI1);

#ifdef __cplusplus
}
#endif
#endif
