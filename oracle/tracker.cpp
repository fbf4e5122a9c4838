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
    L.cr = (r + kFlowRatio - 1) / kFlowRatio;  // domain_div_up (generic_strategy ctor)
    L.cc = (c + kFlowRatio - 1) / kFlowRatio;
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
    L.feats.clear();
    L.matches.clear();
    L.s1 = image<uint8_t>();  // descriptors of set particles are zero until the next frame
    L.s2 = image<uint8_t>();
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
  if (P.profile == 2) return update_level_bc2s(s);
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

// ------------------------------------------------------------- bc2s strategy
// match_particles_kernel (tracking2/tracking_strategies.hpp:274-402) with
// particle_container::before_matching and move / remove applied in ascending
// index order (the sparse buffer keeps the last writer).  Positions and speeds
// are integers (particle::pos is i_short2, the moves are integer); they are
// kept in the float fields.  multiscale_prediction returns
// motion_based_prediction (predictions.h:20-32): no motion for age 1, else the
// speed.  With every descriptor weight 255, cpt = 8 and the test is
// n_diff < 4 over the s2 samples.
void tracker::match_bc2s(int s) {
  level_state& L = levels[s];
  std::fill(L.sparse.px.begin(), L.sparse.px.end(), -1);
  L.compact_has_run = false;
  const bc2s_images F{&L.s1, &L.s2};
  auto has = [&](int r, int c) { return r >= 0 && r < L.nr && c >= 0 && c < L.nc; };
  L.last_decisions.assign(L.parts.size(), lk_decision{0, f2{0.f, 0.f}, 0.f});
  for (size_t i = 0; i < L.parts.size(); i++) {
    particle& p = L.parts[i];
    if (p.age == 0) continue;
    const int r = tz(p.pos_r), c = tz(p.pos_c);
    const int pr = r + (p.age == 1 ? 0 : tz(p.speed_r)), pc = c + (p.age == 1 ? 0 : tz(p.speed_c));
    bool keep = false;
    lk_decision& d = L.last_decisions[i];
    if (has(pr, pc)) {
      const descent_result m = two_step_gradient_descent_match(F, pr, pc, L.feats[i]);
      const int mr = pr + m.dr, mc = pc + m.dc;
      d.dist = m.dist;
      d.tr = f2{(float)(pr - r + m.dr), (float)(pc - c + m.dc)};
      if (has(mr, mc)) {
        const feat16 mf = bc2s_feature(F, mr, mc);
        int n_diff = 0;
        for (int k = 8; k < 16; k++)
          if (std::abs((int)L.feats[i].d[k] - (int)mf.d[k]) > P.k) n_diff++;
        if (n_diff < 4 && !((float)L.contrast(mr, mc) < 3.f)) {
          keep = true;
          // kernel_particle_container::move (particle_container.hpp:422-466)
          p.age++;
          p.speed_r = d.tr.r;
          p.speed_c = d.tr.c;
          p.pos_r = (float)mr;
          p.pos_c = (float)mc;
          if (norml2(p.speed_r, p.speed_c) > 0.f) L.feats[i] = mf;
          L.sparse(mr, mc) = (int32_t)i;
        }
      }
    }
    d.accept = keep ? 1 : 0;
    if (!keep) p.age = 0;
  }
}

// compute_flow_stats_kernel + flow_fusion_kernel / flow_fusion_kernel_root
// (tracking_strategies.hpp:537-634, 688-701): per 8x8 cell, the number of
// pixels holding a particle of age >= 1 and the speed of the last one in
// raster order (the reference assigns, it does not sum); then the fused count
// and flow.  An upper cell bin / 2 outside the coarser scale's cell grid (a
// side of 16k + 1 pixels) is read as empty (count 0, no flow): the reference
// reads past its row there.
void tracker::flow_stats_bc2s(int s) {
  level_state& L = levels[s];
  L.st_count.assign((size_t)L.cr * L.cc, 0);
  L.st_speed.assign((size_t)L.cr * L.cc, f2{0.f, 0.f});
  for (int R = 0; R < L.cr; R++)
    for (int C = 0; C < L.cc; C++) {
      int count = 0;
      f2 last{0.f, 0.f};
      for (int r = 0; r < kFlowRatio; r++)
        for (int c = 0; c < kFlowRatio; c++) {
          const int y = R * kFlowRatio + r, x = C * kFlowRatio + c;
          if (!L.sparse.has(y, x) || L.sparse(y, x) == -1) continue;
          const particle& q = L.parts[L.sparse(y, x)];
          if (q.age >= 1) {
            count++;
            last = f2{q.speed_r, q.speed_c};
          }
        }
      L.st_count[(size_t)R * L.cc + C] = count;
      L.st_speed[(size_t)R * L.cc + C] = last;
    }
  L.flow = image<f2>(L.cr, L.cc, f2{kNoFlow, kNoFlow});
  L.msc = image<uint32_t>(L.cr, L.cc, 0u);
  const bool root = s + 1 >= (int)levels.size();
  for (int R = 0; R < L.cr; R++)
    for (int C = 0; C < L.cc; C++) {
      const int n = L.st_count[(size_t)R * L.cc + C];
      const f2 v = L.st_speed[(size_t)R * L.cc + C];
      if (root) {
        L.msc(R, C) = (uint32_t)n;
        if (n > 1) L.flow(R, C) = f2{v.r / (float)n, v.c / (float)n};
      } else {
        const level_state& U = levels[s + 1];
        const bool in = R / 2 < U.cr && C / 2 < U.cc;
        const int un = in ? U.st_count[(size_t)(R / 2) * U.cc + C / 2] : 0;
        const f2 uf = in ? U.flow(R / 2, C / 2) : f2{kNoFlow, kNoFlow};
        L.msc(R, C) = (uint32_t)(n + un);
        if (n > 1) L.flow(R, C) = v;
        else if (uf.r != kNoFlow || uf.c != kNoFlow) L.flow(R, C) = f2{uf.r * 2.f, uf.c * 2.f};
      }
    }
}

// filter_bad_particles_kernel (tracking_strategies.hpp:491-525) on every
// particle of the dense vector.
void tracker::filter_bc2s(level_state& L) {
  for (particle& p : L.parts) {
    const int R = tz(p.pos_r) / kFlowRatio, C = tz(p.pos_c) / kFlowRatio;
    const uint32_t m = L.msc(R, C);
    if (p.age > 2 && m > 0) {
      const f2 f = L.flow(R, C);
      if (norml2(p.speed_r - f.r, p.speed_c - f.c) > 10.f) {
        p.age = 0;
        continue;
      }
    }
    bool alone = true;
    for (int k = 0; k < 8; k++) {
      const int y = R + kC8[k][0], x = C + kC8[k][1];
      if (L.msc.has(y, x) && L.msc(y, x) != 0) alone = false;
    }
    if (alone && m == 1) p.age = 0;
  }
}

// generic_strategy::update (tracking_strategies.hpp:155-198) with
// bc2s_feature<cpu>::update built without OpenCV (bc2s_feature.hpp:232-253:
// two 1-2-1 blurs, then the borders clamp-filled) and
// generic_strategy::create_detector_mask (:228-272: the particle mask, then 0
// where the contrast is below 5).  estimate_camera_motion sets zero motions
// that nothing reads.
void tracker::update_level_bc2s(int s) {
  level_state& L = levels[s];
  L.cur_u8 = L.in_u8;
  contrast_u32(L.cur_u8, L.contrast);
  blur121_u8(L.cur_u8, L.s1);
  blur121_u8(L.s1, L.s2);
  L.seen = true;
  match_bc2s(s);
  flow_stats_bc2s(s);
  if (!(L.frame_cpt % P.detector_frequency)) {
    detector_mask(L);
    for (size_t i = 0; i < L.mask.px.size(); i++)
      if (L.contrast.px[i] < 5) L.mask.px[i] = 0;
    fast_saliency(L.cur_u8, L.mask, P.fast_n, (int)P.fast_th, L.sal);
    const std::vector<std::pair<int, int>> pts = fast_select(L.sal, L.sparse, L.box_size, P.fast_th);
    const bc2s_images F{&L.s1, &L.s2};
    for (const auto& rc : pts) L.feats.push_back(bc2s_feature(F, rc.first, rc.second));
    container_add(L.parts, L.sparse, pts);
    // compact(): the features follow their particles
    std::vector<feat16> kept;
    kept.reserve(L.feats.size());
    for (size_t i = 0; i < L.parts.size(); i++)
      if (L.parts[i].age != 0) kept.push_back(L.feats[i]);
    L.feats.swap(kept);
    compact(L);
  }
  if (!(L.frame_cpt % P.filtering_frequency)) {
    filter_bc2s(L);
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
