// Oracle: particle container, pyrlk strategy frame step and tracker recursion.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"
#include "oracle_tracker.hpp"

namespace orc {

// c8_h, CPU table order (cuimg/neighb2d_data.h:80-85), (row, col).
static const int kC8[8][2] = {{-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};

static inline int32_t sparse_at(const image<int32_t>& s, int r, int c) {
  return s.has(r, c) ? s(r, c) : -1;  // border of the sparse buffer is -1
}

tracker::tracker(int nr, int nc, const tracker_params& p) : P(p) {
  int r = nr, c = nc;
  for (int l = 0; l < P.nscales; l++) {
    level_state L;
    L.nr = r;
    L.nc = c;
    L.sparse = image<int32_t>(r, c, -1);  // particle_container ctor memset -1
    L.pdomain = inner(r, c, 3);           // pyrlk_cpu ctor: d - border(3)
    L.box_size = P.box_size;
    levels.push_back(std::move(L));
    r = r / 2;  // obox2d / 2 (tracking2/tracker.hpp:18-26)
    c = c / 2;
  }
  std::vector<float> cf(2 * P.blur_K + 1);
  gauss_coefs(P.blur_sigma, P.blur_K, true, cf.data());
  blur_coef = cf;
}

void tracker::clear() {
  // tracker::clear (tracking2/tracker.hpp:95-102): strategy frame counter and
  // container (sparse <- -1, vectors emptied).
  for (auto& L : levels) {
    L.frame_cpt = 0;
    L.seen = false;
    L.parts.clear();
    L.matches.clear();
    std::fill(L.sparse.px.begin(), L.sparse.px.end(), -1);
  }
}

// tracker::update_input + subsample_input (tracking2/tracker.hpp:53-69).
void tracker::update_input(const uint8_t* frame) {
  level_state& L0 = levels[0];
  L0.in_u8 = image<uint8_t>(L0.nr, L0.nc);
  std::memcpy(L0.in_u8.px.data(), frame, (size_t)L0.nr * L0.nc);
  for (size_t l = 1; l < levels.size(); l++) pyr_down_u8(levels[l - 1].in_u8, levels[l].in_u8);
}

static image<float> to_float(const image<uint8_t>& a) {
  image<float> o(a.nr, a.nc);
  for (size_t i = 0; i < a.px.size(); i++) o.px[i] = (float)a.px[i];
  return o;
}

// Builds the matcher's view of one level from the strategy's prev/cur frames.
void tracker::prepare_matcher_images(level_state& L) {
  if (P.profile == 0) {
    L.A = to_float(L.prev_u8);
    L.B = to_float(L.cur_u8);
    sobel_gradient(L.A, true, L.G);
  } else {
    // the GPU's arithmetic (fused taps) goes with its window-sum order
    const bool fused = P.lk.group_lanes != 0;
    gaussian_sep_f32<uint8_t>(L.prev_u8, blur_coef.data(), P.blur_K, L.A, fused);
    gaussian_sep_f32<uint8_t>(L.cur_u8, blur_coef.data(), P.blur_K, L.B, fused);
    sobel_gradient(L.A, false, L.G);
  }
}

// pyrlk_cpu::get_flow_at (tracking2/pyrlk_tracker.hpp:1025-1063) at level j for
// a synthetic age-1 particle at p.  A failed match returns (-1,-1) and is
// accepted when p + (-1,-1) is in the level domain; the reference falls off the
// end of the function when it is not, which is defined here as identity.
f2 tracker::flow_at(int j, f2 p) const {
  const level_state& L = levels[j];
  f2 pred_tr = prediction(j, p, 1, f2{0.f, 0.f});
  f2 pred{p.r + pred_tr.r, p.c + pred_tr.c};
  if (inner(L.nr, L.nc, 3).has(tz(pred.r), tz(pred.c))) {
    lk_level lv{&L.A, &L.B, &L.G};
    match_result m = lk_match(P.lk, p, pred_tr, lv);
    f2 mt{p.r + m.tr.r, p.c + m.tr.c};
    if (tz(mt.r) >= 0 && tz(mt.r) < L.nr && tz(mt.c) >= 0 && tz(mt.c) < L.nc) return m.tr;
    return f2{0.f, 0.f};
  }
  return f2{0.f, 0.f};
}

// recursive_prediction (tracking2/predictions.h:60-74) with
// translation_transform::scale_transform returning 2*t (transformations.h:77-80).
f2 tracker::prediction(int s, f2 pos, int age, f2 speed) const {
  if (s + 1 < (int)levels.size()) {
    f2 half{(float)((double)pos.r / 2.), (float)((double)pos.c / 2.)};
    f2 f = flow_at(s + 1, half);
    return f2{2 * f.r, 2 * f.c};
  }
  if (age == 1) return f2{0.f, 0.f};
  return speed;
}

// pyrlk_match_particles_kernel (tracking2/pyrlk_tracker.hpp:714-768) for every
// particle: decision only (no container writes), so it can run in parallel.
void tracker::lk_decide(int s, std::vector<lk_decision>& out) const {
  const level_state& L = levels[s];
  const int n = (int)L.parts.size();
  out.assign(n, lk_decision{});
#pragma omp parallel for schedule(static, 2)
  for (int i = 0; i < n; i++) {
    const particle& part = L.parts[i];
    lk_decision d;
    d.accept = 0;
    d.tr = f2{0.f, 0.f};
    d.dist = 0.f;
    const box& dom = L.pdomain;
    if (part.age > 0 && dom.has(tz(part.pos_r), tz(part.pos_c))) {
      f2 pos{part.pos_r, part.pos_c};
      f2 pred_tr = prediction(s, pos, part.age, f2{part.speed_r, part.speed_c});
      f2 pred{pos.r + pred_tr.r, pos.c + pred_tr.c};
      if (dom.has(tz(pred.r), tz(pred.c))) {
        lk_level lv{&L.A, &L.B, &L.G};
        match_result m = lk_match(P.lk, pos, pred_tr, lv);
        d.tr = m.tr;
        d.dist = m.dist;
        int mr = tz(pos.r + m.tr.r), mc = tz(pos.c + m.tr.c);
        if (dom.has(mr, mc) && m.dist < (float)P.k &&
            norml2((float)mr - pred.r, (float)mc - pred.c) < 160.0)
          d.accept = 1;
      }
    }
    out[i] = d;
  }
}

// particle_container::before_matching + the kernel's move/remove applied in
// ascending index order, so the sparse buffer keeps the last writer
// (kernel_particle_container::move, tracking2/particle_container.hpp:422-466).
void tracker::match_particles(int s) {
  level_state& L = levels[s];
  std::fill(L.sparse.px.begin(), L.sparse.px.end(), -1);
  L.compact_has_run = false;
  lk_decide(s, L.last_decisions);
  for (size_t i = 0; i < L.parts.size(); i++) {
    particle& p = L.parts[i];
    const lk_decision& d = L.last_decisions[i];
    if (d.accept) {
      p.age++;
      p.speed_r = d.tr.r;
      p.speed_c = d.tr.c;
      p.pos_r = p.pos_r + d.tr.r;
      p.pos_c = p.pos_c + d.tr.c;
      L.sparse(tz(p.pos_r), tz(p.pos_c)) = (int32_t)i;
    } else {
      p.age = 0;
    }
  }
}

// pyrlk_cpu::create_detector_mask + particle_mask_kernel
// (tracking2/pyrlk_tracker.hpp:879-888, tracking_strategies.hpp:201-225): every
// particle of the dense vector, dead ones included, clears its 3x3 block.
image<uint8_t> particle_mask(const std::vector<particle>& parts, int nr, int nc) {
  image<uint8_t> mask(nr, nc, 255);
  for (const particle& p : parts) {
    int r = tz(p.pos_r), c = tz(p.pos_c);
    if (mask.has(r, c)) mask(r, c) = 0;
    for (int k = 0; k < 8; k++)
      if (mask.has(r + kC8[k][0], c + kC8[k][1])) mask(r + kC8[k][0], c + kC8[k][1]) = 0;
  }
  return mask;
}

void tracker::detector_mask(level_state& L) { L.mask = particle_mask(L.parts, L.nr, L.nc); }

// fast_detector::new_particles(cpu) (tracking2/fast_detector.hpp:217-262),
// selection part: one box of box_size^2 pixels at p*box + (1,1) per cell of
// domain/box; a box holding any particle (sparse index != -1) is skipped; the
// first strict maximum in raster order wins; saliency below the float
// threshold is rejected; the points come out in raster order over
// domain - border(3), the order in which new_particles adds them.
std::vector<std::pair<int, int>> fast_select(const image<uint8_t>& sal, const image<int32_t>& sparse, int b,
                                             float th) {
  const int nr = sal.nr, nc = sal.nc;
  const int br = (int)((float)nr / (float)b), bc = (int)((float)nc / (float)b);
  image<uint8_t> np(nr, nc, 0);
#pragma omp parallel for schedule(static, 2)
  for (int i = 0; i < br; i++)
    for (int j = 0; j < bc; j++) {
      int pr = i * b + 1, pc = j * b + 1;
      bool skip = false;
      for (int r = 0; r < b && !skip; r++)
        for (int c = 0; c < b; c++)
          if (sal.has(pr + r, pc + c) && sparse_at(sparse, pr + r, pc + c) != -1) {
            skip = true;
            break;
          }
      if (skip) continue;
      // only box == 1 reaches past the domain here, where the saliency border is 0
      float vmax = sal.has(pr, pc) ? (float)sal(pr, pc) : 0.f;
      int mr = pr, mc = pc;
      for (int r = 0; r < b; r++)
        for (int c = 0; c < b; c++)
          if (sal.has(pr + r, pc + c)) {
            float sv = (float)sal(pr + r, pc + c);
            if (vmax < sv) {
              vmax = sv;
              mr = pr + r;
              mc = pc + c;
            }
          }
      if (vmax < th || !np.has(mr, mc)) continue;
      np(mr, mc) = 1;
    }
  std::vector<std::pair<int, int>> out;
  const box in3 = inner(nr, nc, 3);
  for (int r = in3.r1; r <= in3.r2; r++)
    for (int c = in3.c1; c <= in3.c2; c++)
      if (np(r, c)) out.emplace_back(r, c);
  return out;
}

// particle_container::add (tracking2/particle_container.hpp:370-387) for each
// point in order: a particle of age 1 and zero speed at the point, sparse(point)
// = its index.  Points outside the sparse image's domain write no index.
void container_add(std::vector<particle>& parts, image<int32_t>& sparse,
                   const std::vector<std::pair<int, int>>& pts) {
  for (const auto& rc : pts) {
    particle p{};
    p.age = 1;
    p.pos_r = (float)rc.first;
    p.pos_c = (float)rc.second;
    if (sparse.has(rc.first, rc.second)) sparse(rc.first, rc.second) = (int32_t)parts.size();
    parts.push_back(p);
  }
}

void tracker::new_particles(level_state& L) {
  container_add(L.parts, L.sparse, fast_select(L.sal, L.sparse, L.box_size, P.fast_th));
}

// particle_container::compact(cpu) (tracking2/particle_container.hpp:216-252):
// particles with age != 0 keep their order; matches[old] = new index or
// 0xFFFFFFFF; each kept particle writes its new index at its truncated position
// (the last one of a pixel wins; positions outside the domain write nothing).
void container_compact(std::vector<particle>& parts, image<int32_t>& sparse, std::vector<uint32_t>& matches) {
  matches.assign(parts.size(), 0xFFFFFFFFu);
  size_t j = 0;
  for (size_t i = 0; i < parts.size(); i++) {
    if (parts[i].age != 0) {
      particle p = parts[i];
      parts[j] = p;
      if (sparse.has(tz(p.pos_r), tz(p.pos_c))) sparse(tz(p.pos_r), tz(p.pos_c)) = (int32_t)j;
      matches[i] = (uint32_t)j;
      j++;
    }
  }
  parts.resize(j);
}

void tracker::compact(level_state& L) {
  L.compact_has_run = true;
  container_compact(L.parts, L.sparse, L.matches);
}

// merge_trajectories (tracking2/merge_trajectories.h:10-35) over the dense
// vector in ascending index order: a live particle removes the first c8
// neighbour (CPU table order) that is more than 2 frames older and whose speed
// differs by less than 2 px.
void tracker::merge(level_state& L) {
  for (size_t i = 0; i < L.parts.size(); i++) {
    const particle part = L.parts[i];
    if (part.age == 0) continue;
    int r = tz(part.pos_r), c = tz(part.pos_c);
    for (int k = 0; k < 8; k++) {
      int32_t bi = sparse_at(L.sparse, r + kC8[k][0], c + kC8[k][1]);
      if (bi != -1) {
        particle& buddy = L.parts[bi];
        if (buddy.age > part.age + 2 &&
            norml2(part.speed_r - buddy.speed_r, part.speed_c - buddy.speed_c) < 2.f) {
          buddy.age = 0;
          break;
        }
      }
    }
  }
}

// Standalone merge on a given container state (tests).
void merge_state(std::vector<particle>& parts, const image<int32_t>& sparse) {
  level_state L;
  L.parts.swap(parts);
  L.sparse = sparse;
  L.nr = sparse.nr;
  L.nc = sparse.nc;
  tracker::merge(L);
  parts.swap(L.parts);
}

// pyrlk_cpu::update (tracking2/pyrlk_tracker.hpp:831-875).  The strategy's
// copy of its input is an exact copy (the reference's equal-border memcpy skips
// the last `border` rows, copy.h:35, which leaves them uninitialised).
// filter() is a no-op: its multiscale counts are never set to a non-zero value
// (the fusion kernel is commented out, :958-968) and the finest level's counts
// are uninitialised memory, defined here as zero.
// After a scale's own clear() (frame counter 0, tracker not cleared) the
// reference copies the current frame into prev_input_ and leaves its gradient
// from two frames before, which finer scales then see through get_flow_at;
// here the scale keeps its two most recent frames instead (the matcher images
// of the previous and the current frame), which is what its frame counter
// restarting changes nothing about.
void tracker::update_level(int s) {
  level_state& L = levels[s];
  if (L.frame_cpt > 0 || L.seen) L.prev_u8 = L.cur_u8;
  else L.prev_u8 = L.in_u8;
  L.cur_u8 = L.in_u8;
  if (L.seen) prepare_matcher_images(L);
  L.seen = true;

  if (L.frame_cpt > 0) match_particles(s);
  if (!(L.frame_cpt % P.detector_frequency)) {
    detector_mask(L);
    fast_saliency(L.cur_u8, L.mask, P.fast_n, (int)P.fast_th, L.sal);
    new_particles(L);
    compact(L);
  }
  if (!(L.frame_cpt % P.filtering_frequency)) {
    if (P.with_merge) merge(L);
  }
  L.frame_cpt++;
}

// tracker::run (tracking2/tracker.hpp:73-91): coarsest level first.
void tracker::run(const uint8_t* frame) {
  update_input(frame);
  step();
}

void tracker::step() {
  for (int s = (int)levels.size() - 1; s >= 0; s--) update_level(s);
}

}  // namespace orc
