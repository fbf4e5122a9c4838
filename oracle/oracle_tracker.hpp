// Oracle tracker state.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#pragma once
#include <utility>

#include "oracle_core.hpp"

namespace orc {

struct tracker_params {
  int nscales = 3;
  int profile = 1;  // 0: reference defaults on u8 frames, 1: blurred FP32 frames,
                    // 2: bc2s descriptor + gradient-descent matcher (generic_strategy)
  lk_params lk;
  int fast_n = 9;
  float fast_th = 20.f;
  int box_size = 3;
  int detector_frequency = 1;
  int filtering_frequency = 1;
  int k = 300;
  int with_merge = 1;
  int blur_sigma = 1;
  int blur_K = 3;
};

struct lk_decision {
  int accept;
  f2 tr;
  float dist;
};

struct level_state {
  int nr = 0, nc = 0;
  image<uint8_t> in_u8;            // tracker::input_ of this scale
  image<uint8_t> prev_u8, cur_u8;  // strategy prev_input_ / input_
  image<float> A, B;               // matcher images (previous / current)
  image<f2> G;                     // gradient of A
  image<uint8_t> mask, sal;
  image<int32_t> sparse;
  std::vector<particle> parts;
  std::vector<uint32_t> matches;
  std::vector<lk_decision> last_decisions;
  bool compact_has_run = false;
  int frame_cpt = 0;
  bool seen = false;  // a frame went through this scale since the last tracker clear
  box pdomain{0, 0, -1, -1};  // pyrlk_cpu::particle_domain_
  int box_size = 3;             // fast_detector::set_box_size of this scale
  // bc2s strategy (generic_strategy, profile 2)
  image<uint32_t> contrast;     // contrast_
  image<uint8_t> s1, s2;        // feature_ (bc2s_feature)
  std::vector<feat16> feats;    // pset.features(), parallel to parts
  int cr = 0, cc = 0;           // flow cells: domain_div_up(domain, 8)
  std::vector<int> st_count;    // flow_stats_ .first
  std::vector<f2> st_speed;     // flow_stats_ .second
  image<f2> flow;               // flow_
  image<uint32_t> msc;          // multiscale_count_
};

// NO_FLOW (tracking2/tracker_defines.h:4)
constexpr float kNoFlow = -500.f;
constexpr int kFlowRatio = 8;  // generic_strategy::flow_ratio

// Detector mask of a dense vector (255, 0 on the 3x3 block of every particle).
image<uint8_t> particle_mask(const std::vector<particle>& parts, int nr, int nc);
// Box selection of new points, in the raster order they are added.
std::vector<std::pair<int, int>> fast_select(const image<uint8_t>& sal, const image<int32_t>& sparse, int box,
                                             float th);
// particle_container::add over a point list, and compact(cpu).
void container_add(std::vector<particle>& parts, image<int32_t>& sparse,
                   const std::vector<std::pair<int, int>>& pts);
void container_compact(std::vector<particle>& parts, image<int32_t>& sparse, std::vector<uint32_t>& matches);

struct tracker {
  tracker_params P;
  std::vector<level_state> levels;
  std::vector<float> blur_coef;

  tracker(int nr, int nc, const tracker_params& p);
  void run(const uint8_t* frame);
  void step();  // tracker::run() on the current input
  void clear();

  void update_input(const uint8_t* frame);
  void update_level(int s);
  void prepare_matcher_images(level_state& L);
  f2 flow_at(int j, f2 p) const;
  f2 prediction(int s, f2 pos, int age, f2 speed) const;
  void lk_decide(int s, std::vector<lk_decision>& out) const;
  void match_particles(int s);
  void detector_mask(level_state& L);
  void new_particles(level_state& L);
  void compact(level_state& L);
  static void merge(level_state& L);
  // generic_strategy<bc2s_feature, fast_detector> (profile 2)
  void update_level_bc2s(int s);
  void match_bc2s(int s);
  void flow_stats_bc2s(int s);
  void filter_bc2s(level_state& L);
};

void merge_state(std::vector<particle>& parts, const image<int32_t>& sparse);

}  // namespace orc
