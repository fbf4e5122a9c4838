// C entry points of the oracle for ctypes.  TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

#include <omp.h>

#include "oracle_core.hpp"
#include "oracle_tracker.hpp"

namespace orc {
double lk_loss_grad_f64(const image<double>& I0, const image<double>& I1, const double* coef,
                        int K, const double* pos2, const double* d2, int np, int ws,
                        double* dL_dd2, image<double>* dL_dI1);
}

using namespace orc;

// memcpy of `bytes` bytes; an empty copy may have null pointers (memcpy's may not).
static void copy_bytes(void* dst, const void* src, size_t bytes) {
  if (bytes) std::memcpy(dst, src, bytes);
}

template <typename T>
static image<T> wrap(const T* p, int nr, int nc) {
  image<T> im(nr, nc);
  copy_bytes(im.px.data(), p, sizeof(T) * (size_t)nr * nc);
  return im;
}

extern "C" {

float orc_gauss_coef_runtime(float sigma, int x) { return gauss_coef_runtime(sigma, x); }
float orc_gauss_coef_meta(int sigma, int x) { return gauss_coef_meta(sigma, x); }
void orc_gauss_coefs(int sigma, int K, int meta_rounding, float* out) {
  gauss_coefs(sigma, K, meta_rounding != 0, out);
}

void orc_gaussian_sep_f32(const float* in, int nr, int nc, int ch, const float* coef, int K, int fused,
                          float* out) {
  for (int k = 0; k < ch; k++) {
    image<float> a(nr, nc), o;
    for (size_t i = 0; i < a.px.size(); i++) a.px[i] = in[i * ch + k];
    gaussian_sep_f32<float>(a, coef, K, o, fused != 0);
    for (size_t i = 0; i < o.px.size(); i++) out[i * ch + k] = o.px[i];
  }
}

void orc_gaussian_sep_u8(const uint8_t* in, int nr, int nc, const float* coef, int K, int fused, float* out) {
  image<float> o;
  gaussian_sep_f32<uint8_t>(wrap(in, nr, nc), coef, K, o, fused != 0);
  copy_bytes(out, o.px.data(), sizeof(float) * o.px.size());
}

void orc_gaussian_sep_f64(const double* in, int nr, int nc, int ch, const double* coef, int K,
                          double* out) {
  for (int k = 0; k < ch; k++) {
    image<double> a(nr, nc), o;
    for (size_t i = 0; i < a.px.size(); i++) a.px[i] = in[i * ch + k];
    gaussian_sep_f64(a, coef, K, o);
    for (size_t i = 0; i < o.px.size(); i++) out[i * ch + k] = o.px[i];
  }
}

void orc_pyr_down_u8(const uint8_t* in, int nr, int nc, uint8_t* out) {
  image<uint8_t> o;
  pyr_down_u8(wrap(in, nr, nc), o);
  copy_bytes(out, o.px.data(), o.px.size());
}

void orc_pyr_down_f32(const float* in, int nr, int nc, float* out) {
  image<float> o;
  pyr_down_f32(wrap(in, nr, nc), o);
  copy_bytes(out, o.px.data(), sizeof(float) * o.px.size());
}

uint8_t orc_fast_score(const uint8_t* in, int nr, int nc, int r, int c, int n, int th) {
  return fast_score(wrap(in, nr, nc), r, c, n, th);
}

void orc_fast_saliency(const uint8_t* in, int nr, int nc, const uint8_t* mask, int n, int th,
                       uint8_t* out) {
  image<uint8_t> s;
  fast_saliency(wrap(in, nr, nc), wrap(mask, nr, nc), n, th, s);
  copy_bytes(out, s.px.data(), s.px.size());
}

void orc_sobel_gradient(const float* prev, int nr, int nc, int integer_source, float* out2) {
  image<f2> g;
  sobel_gradient(wrap(prev, nr, nc), integer_source != 0, g);
  copy_bytes(out2, g.px.data(), sizeof(f2) * g.px.size());
}

void orc_lk_match(const float* A, const float* B, const float* G, int nr, int nc, int ws,
                  int max_iters, float eps, int top_n, int template_u8, int sum_lanes, float p_r, float p_c,
                  float pred_r, float pred_c, float* out_tr2, float* out_dist) {
  image<float> a = wrap(A, nr, nc), b = wrap(B, nr, nc);
  image<f2> g = wrap(reinterpret_cast<const f2*>(G), nr, nc);
  lk_params P;
  P.ws = ws;
  P.max_iters = max_iters;
  P.eps = eps;
  P.top_n = top_n;
  P.template_u8 = template_u8;
  P.group_lanes = sum_lanes;
  lk_level L{&a, &b, &g};
  match_result m = lk_match(P, f2{p_r, p_c}, f2{pred_r, pred_c}, L);
  out_tr2[0] = m.tr.r;
  out_tr2[1] = m.tr.c;
  *out_dist = m.dist;
}

static tracker_params to_params(const orc_tracker_params* p) {
  tracker_params t;
  t.nscales = p->nscales;
  t.profile = p->profile;
  t.lk.ws = p->lk_ws;
  t.lk.max_iters = p->lk_max_iters;
  t.lk.eps = p->lk_eps;
  t.lk.top_n = p->lk_top_n;
  t.lk.template_u8 = p->profile == 0 ? 1 : 0;
  t.lk.group_lanes = p->sum_lanes;
  t.fast_n = p->fast_n;
  t.fast_th = p->fast_th;
  t.box_size = p->box_size;
  t.detector_frequency = p->detector_frequency;
  t.filtering_frequency = p->filtering_frequency;
  t.k = p->k;
  t.with_merge = p->with_merge;
  t.blur_sigma = p->blur_sigma;
  t.blur_K = p->blur_K;
  return t;
}

void* orc_tracker_create(int nr, int nc, const orc_tracker_params* p) {
  return new tracker(nr, nc, to_params(p));
}
void orc_tracker_destroy(void* t) { delete static_cast<tracker*>(t); }
void orc_tracker_clear(void* t) { static_cast<tracker*>(t)->clear(); }
void orc_tracker_run(void* t, const uint8_t* frame) { static_cast<tracker*>(t)->run(frame); }
void orc_tracker_update_input(void* t, const uint8_t* frame) { static_cast<tracker*>(t)->update_input(frame); }
void orc_tracker_step(void* t) { static_cast<tracker*>(t)->step(); }
void orc_tracker_clear_scale(void* t, int s) { static_cast<tracker*>(t)->levels[s].frame_cpt = 0; }
void orc_tracker_set_scale(void* t, int s, int box_size, int r1, int c1, int r2, int c2) {
  level_state& L = static_cast<tracker*>(t)->levels[s];
  L.box_size = box_size;
  L.pdomain = box{r1, c1, r2, c2};
}
int orc_tracker_nscales(void* t) { return (int)static_cast<tracker*>(t)->levels.size(); }

void orc_tracker_level_dims(void* t, int s, int* nr, int* nc) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  *nr = L.nr;
  *nc = L.nc;
}

int orc_tracker_nparticles(void* t, int s) {
  return (int)static_cast<tracker*>(t)->levels[s].parts.size();
}

int orc_tracker_get_particles(void* t, int s, orc_particle* out, int cap) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  int n = (int)L.parts.size();
  if (n > cap) return -n;
  copy_bytes(out, L.parts.data(), sizeof(particle) * n);
  return n;
}

// Profile 2: each particle's descriptor is taken at its position in the
// scale's current feature images (zero before the first frame).
int orc_tracker_set_particles(void* t, int s, const orc_particle* in, int n) {
  tracker& T = *static_cast<tracker*>(t);
  level_state& L = T.levels[s];
  L.parts.resize(n);
  copy_bytes(L.parts.data(), in, sizeof(particle) * n);
  if (T.P.profile == 2) {
    L.feats.assign(n, feat16{});
    if (!L.s1.px.empty())
      for (int i = 0; i < n; i++)
        L.feats[i] = bc2s_feature(bc2s_images{&L.s1, &L.s2}, tz(L.parts[i].pos_r), tz(L.parts[i].pos_c));
  }
  return n;
}

int orc_tracker_get_matches(void* t, int s, uint32_t* out, int cap) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  int n = (int)L.matches.size();
  if (n > cap) return -n;
  copy_bytes(out, L.matches.data(), sizeof(uint32_t) * n);
  return n;
}

int orc_tracker_compact_has_run(void* t, int s) {
  return static_cast<tracker*>(t)->levels[s].compact_has_run ? 1 : 0;
}

void orc_tracker_get_sparse(void* t, int s, int32_t* out) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  copy_bytes(out, L.sparse.px.data(), sizeof(int32_t) * L.sparse.px.size());
}

void orc_tracker_get_saliency(void* t, int s, uint8_t* out) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  if (L.sal.px.empty()) std::memset(out, 0, (size_t)L.nr * L.nc);
  else copy_bytes(out, L.sal.px.data(), L.sal.px.size());
}

void orc_tracker_get_level(void* t, int s, uint8_t* out) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  copy_bytes(out, L.in_u8.px.data(), L.in_u8.px.size());
}

void orc_tracker_get_matcher_images(void* t, int s, float* A, float* B, float* G2) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  if (A) copy_bytes(A, L.A.px.data(), sizeof(float) * L.A.px.size());
  if (B) copy_bytes(B, L.B.px.data(), sizeof(float) * L.B.px.size());
  if (G2) copy_bytes(G2, L.G.px.data(), sizeof(f2) * L.G.px.size());
}

static int copy_decisions(const std::vector<lk_decision>& d, int32_t* accept, float* tr2,
                          float* dist, int cap) {
  int n = (int)d.size();
  if (n > cap) return -n;
  for (int i = 0; i < n; i++) {
    accept[i] = d[i].accept;
    tr2[2 * i] = d[i].tr.r;
    tr2[2 * i + 1] = d[i].tr.c;
    dist[i] = d[i].dist;
  }
  return n;
}

int orc_tracker_get_decisions(void* t, int s, int32_t* accept, float* tr2, float* dist, int cap) {
  return copy_decisions(static_cast<tracker*>(t)->levels[s].last_decisions, accept, tr2, dist, cap);
}

int orc_tracker_lk_decide(void* t, int s, int32_t* accept, float* tr2, float* dist, int cap) {
  std::vector<lk_decision> d;
  static_cast<tracker*>(t)->lk_decide(s, d);
  return copy_decisions(d, accept, tr2, dist, cap);
}

void orc_particle_mask(const orc_particle* parts, int n, int nr, int nc, uint8_t* mask) {
  std::vector<particle> v(n);
  if (n) copy_bytes(v.data(), parts, sizeof(particle) * n);
  const image<uint8_t> m = particle_mask(v, nr, nc);
  copy_bytes(mask, m.px.data(), m.px.size());
}

int orc_fast_select(const uint8_t* sal, const int32_t* sparse, int nr, int nc, int box, float th, int32_t* pts,
                    int cap) {
  const std::vector<std::pair<int, int>> v = fast_select(wrap(sal, nr, nc), wrap(sparse, nr, nc), box, th);
  for (size_t k = 0; k < v.size() && (int)k < cap; k++) {
    pts[2 * k] = v[k].first;
    pts[2 * k + 1] = v[k].second;
  }
  return (int)v.size();
}

void orc_blur121_u8(const uint8_t* in, int nr, int nc, uint8_t* out) {
  image<uint8_t> o;
  blur121_u8(wrap(in, nr, nc), o);
  copy_bytes(out, o.px.data(), o.px.size());
}

void orc_gaussian_int_taps(float sigma, int hw, int32_t* out) {
  const std::vector<int> t = int_gauss_taps(sigma, hw);
  for (size_t i = 0; i < t.size(); i++) out[i] = t[i];
}

void orc_gaussian_int_u8(const uint8_t* in, int nr, int nc, float sigma, int hw, uint8_t* out) {
  image<uint8_t> o;
  gaussian_int_u8(wrap(in, nr, nc), sigma, hw, o);
  copy_bytes(out, o.px.data(), o.px.size());
}

void orc_contrast_u32(const uint8_t* in, int nr, int nc, uint32_t* out) {
  image<uint32_t> o;
  contrast_u32(wrap(in, nr, nc), o);
  copy_bytes(out, o.px.data(), sizeof(uint32_t) * o.px.size());
}

void orc_bc2s_feature(const uint8_t* s1, const uint8_t* s2, int nr, int nc, int r, int c, uint8_t* out16) {
  const image<uint8_t> a = wrap(s1, nr, nc), b = wrap(s2, nr, nc);
  const feat16 f = bc2s_feature(bc2s_images{&a, &b}, r, c);
  std::memcpy(out16, f.d, 16);
}

int orc_bc2s_distance(const uint8_t* s1, const uint8_t* s2, int nr, int nc, const uint8_t* f16, int r, int c,
                      int scale) {
  const image<uint8_t> a = wrap(s1, nr, nc), b = wrap(s2, nr, nc);
  feat16 f;
  std::memcpy(f.d, f16, 16);
  return bc2s_distance(bc2s_images{&a, &b}, f, r, c, scale);
}

void orc_bc2s_match(const uint8_t* s1, const uint8_t* s2, int nr, int nc, const uint8_t* f16, int pred_r,
                    int pred_c, int scale, int32_t* out_d2, float* out_dist) {
  const image<uint8_t> a = wrap(s1, nr, nc), b = wrap(s2, nr, nc);
  feat16 f;
  std::memcpy(f.d, f16, 16);
  const bc2s_images F{&a, &b};
  const descent_result m = scale == 0 ? two_step_gradient_descent_match(F, pred_r, pred_c, f)
                                      : gradient_descent_match(F, pred_r, pred_c, f, scale);
  out_d2[0] = m.dr;
  out_d2[1] = m.dc;
  *out_dist = m.dist;
}

int orc_tracker_get_features(void* t, int s, uint8_t* out16, int cap) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  const int n = (int)L.feats.size();
  if (n > cap) return -n;
  copy_bytes(out16, L.feats.data(), 16 * (size_t)n);
  return n;
}

void orc_tracker_get_bc2s_images(void* t, int s, uint8_t* s1, uint8_t* s2, uint32_t* contrast) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  const size_t N = (size_t)L.nr * L.nc;
  auto put = [&](void* dst, const void* src, size_t bytes, size_t have) {
    if (!dst) return;
    if (have) copy_bytes(dst, src, bytes);
    else std::memset(dst, 0, bytes);
  };
  put(s1, L.s1.px.data(), N, L.s1.px.size());
  put(s2, L.s2.px.data(), N, L.s2.px.size());
  put(contrast, L.contrast.px.data(), 4 * N, L.contrast.px.size());
}

void orc_tracker_flow_dims(void* t, int s, int* cr, int* cc) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  *cr = L.cr;
  *cc = L.cc;
}

void orc_tracker_get_flow(void* t, int s, float* flow2, uint32_t* msc) {
  const level_state& L = static_cast<tracker*>(t)->levels[s];
  const size_t N = (size_t)L.cr * L.cc;
  if (flow2) {
    if (L.flow.px.size() == N) copy_bytes(flow2, L.flow.px.data(), sizeof(f2) * N);
    else
      for (size_t i = 0; i < 2 * N; i++) flow2[i] = kNoFlow;
  }
  if (msc) {
    if (L.msc.px.size() == N) copy_bytes(msc, L.msc.px.data(), 4 * N);
    else std::memset(msc, 0, 4 * N);
  }
}

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads() { return omp_get_max_threads(); }

int orc_particles_add(orc_particle* parts, int n, int cap, int32_t* sparse, int nr, int nc, const int32_t* pts,
                      int npts) {
  std::vector<particle> v(n);
  if (n) copy_bytes(v.data(), parts, sizeof(particle) * n);
  image<int32_t> sp = wrap(sparse, nr, nc);
  std::vector<std::pair<int, int>> q;
  for (int k = 0; k < npts && n + k < cap; k++) q.emplace_back(pts[2 * k], pts[2 * k + 1]);
  container_add(v, sp, q);
  if (!v.empty()) copy_bytes(parts, v.data(), sizeof(particle) * v.size());
  copy_bytes(sparse, sp.px.data(), sizeof(int32_t) * sp.px.size());
  return (int)v.size();
}

int orc_particles_compact(orc_particle* parts, int n, int32_t* sparse, int nr, int nc, uint32_t* matches) {
  std::vector<particle> v(n);
  if (n) copy_bytes(v.data(), parts, sizeof(particle) * n);
  image<int32_t> sp = wrap(sparse, nr, nc);
  std::vector<uint32_t> m;
  container_compact(v, sp, m);
  if (!v.empty()) copy_bytes(parts, v.data(), sizeof(particle) * v.size());
  if (n) copy_bytes(matches, m.data(), sizeof(uint32_t) * n);
  copy_bytes(sparse, sp.px.data(), sizeof(int32_t) * sp.px.size());
  return (int)v.size();
}

void orc_merge(orc_particle* parts, int n, const int32_t* sparse, int nr, int nc) {
  if (n <= 0) return;
  std::vector<particle> v(n);
  copy_bytes(v.data(), parts, sizeof(particle) * n);
  merge_state(v, wrap(sparse, nr, nc));
  copy_bytes(parts, v.data(), sizeof(particle) * n);
}

double orc_lk_loss_grad_f64_d(const double* I0, const double* I1, int nr, int nc,
                              const double* coef, int K, const double* pos2, const double* d2,
                              int np, int ws, double* dL_dd2, double* dL_dI1) {
  image<double> a = wrap(I0, nr, nc), b = wrap(I1, nr, nc), g;
  double L = lk_loss_grad_f64(a, b, coef, K, pos2, d2, np, ws, dL_dd2, dL_dI1 ? &g : nullptr);
  if (dL_dI1) copy_bytes(dL_dI1, g.px.data(), sizeof(double) * g.px.size());
  return L;
}

double orc_lk_loss_grad_f64(const float* I0, const float* I1, int nr, int nc, const float* coef,
                            int K, const float* pos2, const float* d2, int np, int ws,
                            double* dL_dd2, double* dL_dI1) {
  std::vector<double> a(I0, I0 + (size_t)nr * nc), b(I1, I1 + (size_t)nr * nc);
  std::vector<double> cf(coef, coef + 2 * K + 1), p(pos2, pos2 + 2 * np), d(d2, d2 + 2 * np);
  return orc_lk_loss_grad_f64_d(a.data(), b.data(), nr, nc, cf.data(), K, p.data(), d.data(), np,
                                ws, dL_dd2, dL_dI1);
}

}  // extern "C"
