// Device-side state of the pyramid LK tracker, shared by particles.hip,
// lk.hip and tracker.hip (and the device headers *_dev.hpp).
#pragma once

#include "common.hpp"

#include <cmath>

namespace cuimg {

constexpr int kMaxLevels = 8;

// Per-scale counters in device memory: particle counts and the frame-step
// state live on the device, so a frame step is a fixed launch sequence with no
// host round trip (and can be replayed as a graph).
struct level_counters {
  uint32_t n;        // dense_particles().size()
  uint32_t n_old;    // size before the last compaction (= matches().size() - n_new)
  uint32_t n_live;   // survivors of the last compaction
  uint32_t n_new;    // particles added by the last detection
  uint32_t overflow; // capacity exceeded (particles dropped)
  uint32_t actors;   // merge: particles with an eligible neighbour
  uint32_t n_matched; // size of the dense vector seen by the last matcher launch
  uint32_t frame;    // frame steps this scale has completed
  uint32_t cur;      // index (0/1) of the current dense vector
  uint32_t compact_has_run;  // the last frame step ran the detector
  uint32_t pend1;    // merge: actors left open by round 1 (in merge_args::pend)
  uint32_t steps;    // scale 0 only: frame steps of the stream since clear (its
                     // parity picks the frame slot of every scale)
  uint32_t merge_restore;  // merge actors of the last frame step whose scratch
                           // entries the next step's fill resets
  uint32_t reserved_;
  unsigned long long live_total;      // particles (age > 0) run through the matcher
  unsigned long long accepted_total;  // of which accepted (moved)
};

// One pyramid level as the matcher sees it: A = previous frame, B = current,
// G = Sobel/8 gradient of A (FP32 frames in a frame step; NULL: the matcher
// computes it from A itself).
struct lk_level_view {
  const void* A;
  const void* B;
  int pitchA, pitchB;
  int nr, nc;
  const float2* G;
  int pitchG, pad;
};

struct lk_params_dev {
  int ws, max_iters, top_n;
  float eps, min_eig;
  int k;
  int template_u8;
  float step_sq;  // sqrt_ge_threshold(eps): |delta| >= eps  <=>  |delta|^2 >= step_sq
  float far_sq;   // sqrt_le_bound(4 ws): |pred - v| > 4 ws  <=>  |pred - v|^2 > far_sq
};

// Squared thresholds equivalent to the matcher's norm tests (host code).  sqrtf
// is correctly rounded and monotone, so for s >= 0 (or NaN):
//   sqrtf(s) >= e  <=>  s >= sqrt_ge_threshold(e), the least float whose root is >= e;
//   sqrtf(s) >  L  <=>  s >  sqrt_le_bound(L), the greatest float whose root is <= L.
inline float sqrt_ge_threshold(float e) {
  if (!(e > 0.f)) return e == e ? 0.f : e;  // e <= 0: every s >= 0 passes; NaN: none
  float y = e * e;
  while (y > 0.f && std::sqrt(y) >= e) y = std::nextafter(y, -HUGE_VALF);
  while (std::sqrt(y) < e) y = std::nextafter(y, HUGE_VALF);
  return y;
}
inline float sqrt_le_bound(float L) {
  if (!(L >= 0.f)) return L == L ? -1.f : L;  // L < 0: every s >= 0 passes; NaN: none
  float y = L * L;
  while (y > 0.f && std::sqrt(y) > L) y = std::nextafter(y, -HUGE_VALF);
  while (std::sqrt(std::nextafter(y, HUGE_VALF)) <= L) y = std::nextafter(y, HUGE_VALF);
  return y;
}
inline lk_params_dev make_lk_params(int ws, int max_iters, int top_n, float eps, int k, int template_u8) {
  return lk_params_dev{ws, max_iters, top_n, eps, 1e-4f, k, template_u8, sqrt_ge_threshold(eps),
                       sqrt_le_bound((float)(ws * 4))};
}

struct lk_args {
  lk_level_view lv[kMaxLevels];
  int nlevels;
  int s;                        // scale of the particles
  cuimg_particle_f* parts;
  level_counters* cnt;
  uint32_t* sparse;             // pixel (0,0) of the scale's sparse image
  int sparse_pitch;
  int32_t* dec_accept;
  float2* dec_tr;
  float* dec_dist;
  int pd_r1, pd_c1, pd_r2, pd_c2;  // particle domain (inclusive box)
  lk_params_dev P;
};

// The bc2s feature images of one scale (u8, borders of 3 and 6, clamp-filled).
struct bc2s_view {
  const uint8_t* s1;
  const uint8_t* s2;
  int p1, p2;  // pitches
  int nr, nc;
};

struct detect_args {
  // level geometry
  int nr, nc;
  int box;                      // detector box size
  int nbr, nbc;                 // boxes per column / row
  float fast_th;
  // images
  const uint8_t* sal; int sal_pitch;
  uint32_t* sparse; int sparse_pitch;
  // particles
  cuimg_particle_f* parts;      // current dense vector
  cuimg_particle_f* parts_out;  // compaction target
  uint32_t* matches;
  uint8_t* flags;               // per particle: bit0 alive, bit1 owns its sparse pixel
  uint32_t* blockcnt;           // per 1024-particle block
  uint8_t* boxcode;             // per detector box: 0xFF none, else dr*box+dc
  uint32_t* subcnt;             // per (box row, dr) sub-row
  level_counters* cnt;
  uint32_t cap;
  int nblocks;                  // compaction blocks covering cap
  // bc2s strategy: 16-byte descriptors follow their particles through the
  // compaction (feat -> feat_out); new particles take theirs from the
  // scale's per-pixel descriptor image
  uint4* feat;                  // nullptr: no descriptors
  uint4* feat_out;
  const uint4* desc;
  int desc_pitch;               // in uint4
};

// A particle with at least one eligible merge target: b[k] is the c8
// neighbour k (valid where bit k of mask is set).
struct merge_actor {
  int32_t i;
  uint32_t mask;
  int32_t b[8];
};

struct merge_args {
  int nr, nc;
  cuimg_particle_f* parts;
  const uint32_t* sparse; int sparse_pitch;
  level_counters* cnt;
  merge_actor* actors;       // cap
  uint32_t* zeroer;          // cap; 0xFFFFFFFF between calls
  uint32_t* minunres;        // cap; 0xFFFFFFFF between calls
  unsigned long long* pend;  // pend_cap: open actors beyond the LDS list
  uint32_t cap, pend_cap;    // pend_cap: power of two >= cap; cap < 2^24
  int defer_restore;         // 1: the resolve leaves the scratch reset to the
                             // next frame step (merge_restore_body)
};
constexpr uint32_t kMergeMaxCap = (1u << 24) - 1;
constexpr uint32_t kMaxParticles = 1u << 30;  // standalone container calls

constexpr int kCompactBlock = 1024;  // particles per compaction block
constexpr int kMaxBox = 15;          // largest detector box

// Host helpers of particles.hip.
int launch_fill_u32(cuimg_hip_ctx* ctx, uint32_t* begin, int pitch, int r0, int r1, int c0, int c1,
                    uint32_t v);
// Bytes of merge scratch for a capacity (pend, actors, zeroer, minunres).
size_t merge_scratch_bytes(uint32_t cap);
merge_args merge_scratch_carve(void* base, uint32_t cap);
int launch_merge_reset(cuimg_hip_ctx* ctx, const merge_args& a);
// Scan, round 1 and resolution of one merge (standalone launches).
int launch_merge(cuimg_hip_ctx* ctx, const merge_args& a);

}  // namespace cuimg
