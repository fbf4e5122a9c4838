// CPU oracle for the cuimg tracking hot path — TEST INFRASTRUCTURE ONLY.
//
// This is a plain C++ restatement of the reference algorithms, used by tests/,
// by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// Product code never links or calls it.
//
// Semantics conventions (all follow the reference, citations are relative to the
// reference tree):
//  * every 2-vector is (row, col); gradient component [0] is d/drow
//    (cuimg/improved_builtin.hpp r()/c(), tracking2/pyrlk_tracker.hpp:912-926);
//  * float -> int coordinate conversion truncates toward zero
//    (cuimg/improved_builtin.hpp:343-347, C cast in assign_explicit);
//  * images are read with clamp-to-edge outside their domain.  Inside the
//    reference's allocated borders this is exactly what fill_border_clamp stored
//    (cuimg/border.h:97-137); beyond them the reference reads unrelated memory,
//    which the oracle defines as clamp-to-edge as well.
//
// Build: -O3 -ffp-contract=off (no FMA contraction), so FP32 results are the
// IEEE results of the written operation order.
#pragma once

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <algorithm>

namespace orc {

// Dense row-major image without border; out-of-domain reads clamp.
template <typename T>
struct image {
  int nr = 0, nc = 0;
  std::vector<T> px;
  image() {}
  image(int r, int c, T v = T()) : nr(r), nc(c), px((size_t)r * c, v) {}
  T& operator()(int r, int c) { return px[(size_t)r * nc + c]; }
  const T& operator()(int r, int c) const { return px[(size_t)r * nc + c]; }
  bool has(int r, int c) const { return r >= 0 && r < nr && c >= 0 && c < nc; }
  const T& clamped(int r, int c) const {
    r = r < 0 ? 0 : (r >= nr ? nr - 1 : r);
    c = c < 0 ? 0 : (c >= nc ? nc - 1 : c);
    return px[(size_t)r * nc + c];
  }
};

struct f2 { float r, c; };

// box2d with inclusive corners; `domain - border(b)` (cuimg/border.h:29-35).
struct box {
  int r1, c1, r2, c2;
  bool has(int r, int c) const { return r >= r1 && r <= r2 && c >= c1 && c <= c2; }
};
inline box inner(int nr, int nc, int b) { return box{b, b, nr - b - 1, nc - b - 1}; }

// float -> int toward zero (cuimg/improved_builtin.hpp:343-347).
inline int tz(float v) { return (int)v; }

// norml2: FP32 sum of squares then sqrt (cuimg/builtin_math.h:123-130).
inline float norml2(float a, float b) {
  float r = 0.f;
  r += a * a;
  r += b * b;
  return std::sqrt(r);
}

// particle_f<translation_transform> (tracking2/particle_container.h:45-57):
// {speed(row,col), pos(row,col), unsigned short age} -> 20 bytes with padding.
struct particle {
  float speed_r, speed_c;
  float pos_r, pos_c;
  uint16_t age;
  uint16_t pad;
};
static_assert(sizeof(particle) == 20, "particle_f layout is 20 bytes");

// ---------------------------------------------------------------- gaussian
float gauss_coef_runtime(float sigma, int x);
float gauss_coef_meta(int sigma, int x);
void gauss_coefs(int sigma, int K, bool meta_rounding, float* out);

template <typename TI>
void gaussian_sep_f32(const image<TI>& in, const float* coef, int K, image<float>& out, bool fused = false);
void gaussian_sep_f64(const image<double>& in, const double* coef, int K, image<double>& out);

// ---------------------------------------------------------------- pyramid
void pyr_down_u8(const image<uint8_t>& in, image<uint8_t>& out);
void pyr_down_f32(const image<float>& in, image<float>& out);

// ---------------------------------------------------------------- FAST
uint8_t fast_score(const image<uint8_t>& in, int r, int c, int n, int th);
void fast_saliency(const image<uint8_t>& in, const image<uint8_t>& mask, int n, int th,
                   image<uint8_t>& sal);

// ---------------------------------------------------------------- LK
struct lk_params {
  int ws = 7;           // window side
  int max_iters = 5;    // Gauss-Newton steps
  float eps = 0.1f;     // stop when |delta| < eps
  int top_n = 49;       // strongest-gradient pixels used; >= ws*ws means all
  int template_u8 = 0;  // 1: template samples truncated to uint8 (u8 images)
  float min_eig = 1e-4f;
  // Order of the window sums (G, b, error).  0: raster order, one FP32 add
  // per pixel.  16 or 64: the order of a group of that many lanes, each lane
  // l summing pixels l, l + lanes, ... with fused multiply-adds, the lane
  // sums then added pairwise (see lk.cpp).
  int group_lanes = 0;
};

// One pyramid level as the matcher sees it: A = previous frame, B = current
// frame, G = gradient of A (Sobel/8, zero on the 1-pixel outer ring).
struct lk_level {
  const image<float>* A;
  const image<float>* B;
  const image<f2>* G;
};

struct match_result { f2 tr; float dist; };
match_result lk_match(const lk_params& P, f2 p, f2 pred_tr, const lk_level& L);
void sobel_gradient(const image<float>& prev, bool integer_source, image<f2>& g);

// ---------------------------------------------------------------- bc2s
// Integer 1-2-1 blur (rows then columns, each truncated by / 4), clamp-to-edge.
void blur121_u8(const image<uint8_t>& in, image<uint8_t>& out);
// Integer taps (int)(100 * g(sigma, i)) of gaussian_kernel<cuda_gpu>.
std::vector<int> int_gauss_taps(float sigma, int hw);
// Separable integer Gaussian with those taps (columns then rows), each pass
// truncated by the tap sum into u8, clamp-to-edge.
void gaussian_int_u8(const image<uint8_t>& in, float sigma, int hw, image<uint8_t>& out);
// contrast_kernel: |differences| of the 4 opposite pairs at distance 2.
void contrast_u32(const image<uint8_t>& in, image<uint32_t>& out);

// bc2s descriptor: 8 samples of s1 on the radius-3 ring (every other point)
// and 8 of s2 at twice the offsets.  The weights of the reference's
// descriptor are 255 for every descriptor it builds, so they are not stored.
struct feat16 {
  uint8_t d[16];
};
struct bc2s_images {
  const image<uint8_t>* s1;
  const image<uint8_t>* s2;
};
feat16 bc2s_feature(const bc2s_images& F, int r, int c);
int bc2s_distance(const bc2s_images& F, const feat16& a, int r, int c, int scale);
struct descent_result {
  int dr, dc;
  float dist;
};
descent_result gradient_descent_match(const bc2s_images& F, int pr, int pc, const feat16& f, int scale);
descent_result two_step_gradient_descent_match(const bc2s_images& F, int pr, int pc, const feat16& f);

}  // namespace orc
