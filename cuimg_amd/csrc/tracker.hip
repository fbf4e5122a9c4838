// tracker<pyrlk> frame step on the GPU: per-scale state, orchestration of the
// kernels and the C ABI of the tracker.
//
// Reference call stack (tracking2/): tracker::run (tracker.hpp:73-91) builds the
// input pyramid (update_input / subsample_input, :53-69) and calls the
// strategy's update on every scale, coarsest first; pyrlk_cpu::update
// (pyrlk_tracker.hpp:831-875) matches the particles (frame > 0), runs the
// detector and inserts new particles (every detector_frequency frames, with a
// compaction), then merges trajectories (every filtering_frequency frames).
//
// GPU layout per scale: two u8 pyramid slots (current / previous frame, border
// 3, clamp-filled), two FP32 blurred slots for the blurred-frame profile, the
// detector's mask and saliency, the u32 sparse index image (border 7), two
// dense particle vectors (compaction ping-pongs between them), the matches
// vector, and per-particle scratch for compaction and the merge.  Particle
// counts stay in device memory (level_counters), so a frame step is a fixed
// sequence of launches with no host synchronisation.
#include "tracker_state.hpp"

#include <cstring>
#include <vector>

namespace cuimg {

struct level_res {
  int nr = 0, nc = 0;
  cuimg_img u8[2]{};
  cuimg_img blur[2]{};
  cuimg_img sal{}, mask{}, sparse{};
  uint8_t* boxcode = nullptr;
  uint32_t* subcnt = nullptr;
  cuimg_particle_f* parts[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* matches = nullptr;
  uint8_t* flags = nullptr;
  uint32_t* blockcnt = nullptr;
  int32_t* dec_accept = nullptr;
  float2* dec_tr = nullptr;
  float* dec_dist = nullptr;
  merge_args merge{};  // merge scratch (one allocation)
  uint32_t cap = 0;
  int nblocks = 0, nbr = 0, nbc = 0;
  bool compact_has_run = false;
  cuimg_scale_params sp{};
};

}  // namespace cuimg

struct cuimg_hip_tracker {
  cuimg_hip_ctx* ctx = nullptr;
  cuimg_tracker_params P{};
  int L = 0;
  cuimg::level_res lv[cuimg::kMaxLevels];
  cuimg::level_counters* cnt = nullptr;  // device, one per scale
  cuimg_img frame_buf{};
  int frame_cpt = 0;
  float coef[33]{};
  std::vector<void*> allocs;
};

namespace cuimg {

static int dev_alloc(cuimg_hip_tracker* t, size_t bytes, void** out) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess)
    return set_error(t->ctx, CUIMG_E_NOMEM, "tracker: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  t->allocs.push_back(p);
  *out = p;
  return 0;
}

template <typename T>
static int dev_alloc_n(cuimg_hip_tracker* t, size_t n, T** out) {
  void* p = nullptr;
  CUIMG_TRY(dev_alloc(t, n * sizeof(T), &p));
  *out = static_cast<T*>(p);
  return 0;
}

static void destroy(cuimg_hip_tracker* t) {
  if (!t) return;
  (void)hipStreamSynchronize(t->ctx->stream);
  for (int s = 0; s < t->L; s++) {
    level_res& R = t->lv[s];
    for (int k = 0; k < 2; k++) {
      image_free(t->ctx, &R.u8[k]);
      image_free(t->ctx, &R.blur[k]);
    }
    image_free(t->ctx, &R.sal);
    image_free(t->ctx, &R.mask);
    image_free(t->ctx, &R.sparse);
  }
  image_free(t->ctx, &t->frame_buf);
  for (void* p : t->allocs) (void)hipFree(p);
  delete t;
}

static int fill_sparse(cuimg_hip_tracker* t, level_res& R) {
  const int b = R.sparse.border;
  return launch_fill_u32(t->ctx, static_cast<uint32_t*>(R.sparse.begin), R.sparse.pitch, -b, R.nr + b, -b,
                         R.nc + b, 0xFFFFFFFFu);
}

static int create(cuimg_hip_ctx* ctx, int nrows, int ncols, const cuimg_tracker_params* P,
                  cuimg_hip_tracker** out) {
  if (!ctx || !P || !out) return CUIMG_E_INVALID;
  *out = nullptr;
  if (P->nscales < 1 || P->nscales > 7)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: nscales %d not in 1..7", P->nscales);
  if (P->lk_ws < 1 || P->lk_ws > 15 || !(P->lk_ws & 1))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: LK window %d must be odd and <= 15", P->lk_ws);
  if (P->box_size < 1 || P->box_size > 15)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: detector box %d not in 1..15", P->box_size);
  if (P->detector_frequency < 1 || P->filtering_frequency < 1 || P->fast_th < 0.f || P->fast_n < 1)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: bad detector / filtering settings");
  if (P->profile != 0 && P->profile != 1)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: profile %d not in {0,1}", P->profile);
  if (P->profile == 1 && (P->blur_K < 1 || P->blur_K > 16 || P->blur_sigma < 1))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: blur sigma %d / half width %d", P->blur_sigma, P->blur_K);
  int r = nrows, c = ncols;
  for (int s = 0; s < P->nscales; s++) {
    if (r < 1 || c < 1)
      return set_error(ctx, CUIMG_E_INVALID, "tracker: scale %d of %dx%d is empty", s, nrows, ncols);
    r /= 2;
    c /= 2;
  }
  auto* t = new cuimg_hip_tracker();
  t->ctx = ctx;
  t->P = *P;
  t->L = P->nscales;
  if (P->profile == 1) gaussian_coefs(P->blur_sigma, P->blur_K, t->coef);
  int rc = 0;
  auto fail = [&](int code) {
    destroy(t);
    return code;
  };
  if ((rc = dev_alloc_n(t, t->L, &t->cnt))) return fail(rc);
  if (hipMemsetAsync(t->cnt, 0, sizeof(level_counters) * t->L, ctx->stream) != hipSuccess) return fail(CUIMG_E_NOMEM);
  if ((rc = image_alloc(ctx, nrows, ncols, 0, CUIMG_U8, &t->frame_buf))) return fail(rc);
  r = nrows;
  c = ncols;
  for (int s = 0; s < t->L; s++) {
    level_res& R = t->lv[s];
    R.nr = r;
    R.nc = c;
    R.sp = cuimg_scale_params{P->fast_n, P->fast_th, P->detector_frequency, P->filtering_frequency, P->k,
                              P->with_merge};
    for (int k = 0; k < 2; k++) {
      if ((rc = image_alloc(ctx, r, c, 3, CUIMG_U8, &R.u8[k]))) return fail(rc);
      if (P->profile == 1 && (rc = image_alloc(ctx, r, c, 0, CUIMG_F32, &R.blur[k]))) return fail(rc);
    }
    if ((rc = image_alloc(ctx, r, c, 1, CUIMG_U8, &R.sal))) return fail(rc);
    if ((rc = image_alloc(ctx, r, c, 1, CUIMG_U8, &R.mask))) return fail(rc);
    if ((rc = image_alloc(ctx, r, c, 7, CUIMG_U32, &R.sparse))) return fail(rc);
    R.nbr = (int)((float)r / (float)P->box_size);
    R.nbc = (int)((float)c / (float)P->box_size);
    const size_t nboxes = (size_t)R.nbr * R.nbc;
    R.cap = (uint32_t)std::max<size_t>((size_t)r * c / 4, 4096);
    if (R.cap > kMergeMaxCap)
      return fail(set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: %dx%d exceeds the particle capacity limit", r, c));
    R.nblocks = ceil_div((int)R.cap, 1024);
    if ((rc = dev_alloc_n(t, std::max<size_t>(nboxes, 1), &R.boxcode))) return fail(rc);
    if ((rc = dev_alloc_n(t, std::max<size_t>((size_t)R.nbr * P->box_size, 1), &R.subcnt))) return fail(rc);
    for (int k = 0; k < 2; k++)
      if ((rc = dev_alloc_n(t, R.cap, &R.parts[k]))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.cap + nboxes, &R.matches))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.cap, &R.flags))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.nblocks, &R.blockcnt))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.cap, &R.dec_accept))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.cap, &R.dec_tr))) return fail(rc);
    if ((rc = dev_alloc_n(t, R.cap, &R.dec_dist))) return fail(rc);
    {
      void* m = nullptr;
      if ((rc = dev_alloc(t, merge_scratch_bytes(R.cap), &m))) return fail(rc);
      R.merge = merge_scratch_carve(m, R.cap);
      if ((rc = launch_merge_reset(ctx, R.merge))) return fail(rc);
    }
    // particle_container ctor: sparse <- -1 (tracking2/particle_container.hpp:23-31)
    if ((rc = fill_sparse(t, R))) return fail(rc);
    r /= 2;
    c /= 2;
  }
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(CUIMG_E_NOMEM);
  *out = t;
  return 0;
}

static lk_level_view matcher_view(const cuimg_hip_tracker* t, int s, int prev, int cur) {
  const level_res& R = t->lv[s];
  const cuimg_img& A = t->P.profile == 0 ? R.u8[prev] : R.blur[prev];
  const cuimg_img& B = t->P.profile == 0 ? R.u8[cur] : R.blur[cur];
  return lk_level_view{A.begin, B.begin, A.pitch, B.pitch, R.nr, R.nc};
}

static int update_level(cuimg_hip_tracker* t, int s, int slot) {
  cuimg_hip_ctx* ctx = t->ctx;
  level_res& R = t->lv[s];
  const cuimg_tracker_params& P = t->P;
  uint32_t* sparse = static_cast<uint32_t*>(R.sparse.begin);

  if (t->frame_cpt > 0) {
    // particle_container::before_matching: sparse <- -1
    CUIMG_TRY(fill_sparse(t, R));
    R.compact_has_run = false;
    lk_args a{};
    for (int j = 0; j < t->L; j++) a.lv[j] = matcher_view(t, j, slot ^ 1, slot);
    a.nlevels = t->L;
    a.s = s;
    a.parts = R.parts[R.cur];
    a.cnt = t->cnt + s;
    a.sparse = sparse;
    a.sparse_pitch = R.sparse.pitch;
    a.dec_accept = R.dec_accept;
    a.dec_tr = R.dec_tr;
    a.dec_dist = R.dec_dist;
    // particle_domain_ = domain - border(3) (pyrlk_tracker.hpp:792)
    a.pd_r1 = 3;
    a.pd_c1 = 3;
    a.pd_r2 = R.nr - 4;
    a.pd_c2 = R.nc - 4;
    a.P = lk_params_dev{P.lk_ws, P.lk_max_iters, P.lk_top_n, P.lk_eps, 1e-4f, R.sp.k, P.profile == 0 ? 1 : 0};
    CUIMG_TRY(launch_lk(ctx, a, P.profile == 0));
  }

  if (!(t->frame_cpt % R.sp.detector_frequency)) {
    detect_args d{};
    d.nr = R.nr;
    d.nc = R.nc;
    d.box = P.box_size;
    d.nbr = R.nbr;
    d.nbc = R.nbc;
    d.fast_th = R.sp.fast_th;
    d.sal = static_cast<const uint8_t*>(R.sal.begin);
    d.sal_pitch = R.sal.pitch;
    d.mask = static_cast<uint8_t*>(R.mask.begin);
    d.mask_pitch = R.mask.pitch;
    d.sparse = sparse;
    d.sparse_pitch = R.sparse.pitch;
    d.parts = R.parts[R.cur];
    d.parts_out = R.parts[R.cur ^ 1];
    d.matches = R.matches;
    d.flags = R.flags;
    d.blockcnt = R.blockcnt;
    d.boxcode = R.boxcode;
    d.subcnt = R.subcnt;
    d.cnt = t->cnt + s;
    d.cap = R.cap;
    d.nblocks = R.nblocks;
    CUIMG_TRY(launch_particle_mask(ctx, d));
    CUIMG_TRY(fast_saliency(ctx, &R.u8[slot], &R.mask, &R.sal, R.sp.fast_n, (int)R.sp.fast_th));
    CUIMG_TRY(launch_detect_and_compact(ctx, d));
    R.cur ^= 1;
    R.compact_has_run = true;
  }

  if (!(t->frame_cpt % R.sp.filtering_frequency) && R.sp.with_merge) {
    merge_args m = R.merge;
    m.nr = R.nr;
    m.nc = R.nc;
    m.parts = R.parts[R.cur];
    m.sparse = sparse;
    m.sparse_pitch = R.sparse.pitch;
    m.cnt = t->cnt + s;
    CUIMG_TRY(launch_merge(ctx, m));
  }
  return 0;
}

static int run(cuimg_hip_tracker* t, const cuimg_img* frame) {
  if (!t || !frame) return CUIMG_E_INVALID;
  const level_res& R0 = t->lv[0];
  if (frame->dtype != CUIMG_U8 || frame->nrows != R0.nr || frame->ncols != R0.nc || !frame->begin)
    return set_error(t->ctx, CUIMG_E_INVALID, "tracker_run: frame must be U8 %dx%d", R0.nr, R0.nc);
  const int slot = t->frame_cpt & 1;
  cuimg_img levels[kMaxLevels];
  for (int s = 0; s < t->L; s++) levels[s] = t->lv[s].u8[slot];
  CUIMG_TRY(pyramid_build(t->ctx, frame, levels, t->L));
  if (t->P.profile == 1) {
    cuimg_img outs[kMaxLevels];
    for (int s = 0; s < t->L; s++) outs[s] = t->lv[s].blur[slot];
    CUIMG_TRY(convolve_sep_multi(t->ctx, levels, outs, t->L, t->coef, t->P.blur_K, 0));
  }
  for (int s = t->L - 1; s >= 0; s--) CUIMG_TRY(update_level(t, s, slot));
  t->frame_cpt++;
  return 0;
}

static int read_counters(cuimg_hip_tracker* t, int s, level_counters* c) {
  CUIMG_HIP(t->ctx, hipMemcpyAsync(c, t->cnt + s, sizeof(level_counters), hipMemcpyDeviceToHost, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

static int check_scale(cuimg_hip_tracker* t, int s) {
  if (!t) return CUIMG_E_INVALID;
  if (s < 0 || s >= t->L) return set_error(t->ctx, CUIMG_E_INVALID, "tracker: scale %d not in 0..%d", s, t->L - 1);
  return 0;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_tracker_create(cuimg_hip_ctx* ctx, int nrows, int ncols, const cuimg_tracker_params* params,
                             cuimg_hip_tracker** out) {
  return create(ctx, nrows, ncols, params, out);
}

int cuimg_hip_tracker_destroy(cuimg_hip_tracker* t) {
  destroy(t);
  return 0;
}

int cuimg_hip_tracker_run(cuimg_hip_tracker* t, const cuimg_img* frame) { return run(t, frame); }

int cuimg_hip_tracker_run_host(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch) {
  if (!t || !frame) return CUIMG_E_INVALID;
  CUIMG_HIP(t->ctx, hipMemcpy2DAsync(t->frame_buf.begin, t->frame_buf.pitch, frame, pitch, t->frame_buf.ncols,
                                     t->frame_buf.nrows, hipMemcpyHostToDevice, t->ctx->stream));
  return run(t, &t->frame_buf);
}

int cuimg_hip_tracker_set_scale_params(cuimg_hip_tracker* t, int scale, const cuimg_scale_params* in) {
  CUIMG_TRY(check_scale(t, scale));
  if (!in || in->fast_n < 1 || in->fast_th < 0.f || in->detector_frequency < 1 || in->filtering_frequency < 1)
    return set_error(t->ctx, CUIMG_E_INVALID, "set_scale_params: bad settings");
  t->lv[scale].sp = *in;
  return 0;
}

int cuimg_hip_tracker_get_scale_params(cuimg_hip_tracker* t, int scale, cuimg_scale_params* out) {
  CUIMG_TRY(check_scale(t, scale));
  if (!out) return CUIMG_E_INVALID;
  *out = t->lv[scale].sp;
  return 0;
}

int cuimg_hip_tracker_clear(cuimg_hip_tracker* t) {
  if (!t) return CUIMG_E_INVALID;
  t->frame_cpt = 0;
  for (int s = 0; s < t->L; s++) {
    CUIMG_TRY(fill_sparse(t, t->lv[s]));
    CUIMG_TRY(launch_merge_reset(t->ctx, t->lv[s].merge));
    t->lv[s].compact_has_run = false;
  }
  CUIMG_HIP(t->ctx, hipMemsetAsync(t->cnt, 0, sizeof(level_counters) * t->L, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_nparticles(cuimg_hip_tracker* t, int scale, uint32_t* n) {
  CUIMG_TRY(check_scale(t, scale));
  level_counters c;
  CUIMG_TRY(read_counters(t, scale, &c));
  if (n) *n = c.n;
  if (c.overflow)
    return set_error(t->ctx, CUIMG_E_CAPACITY, "tracker: scale %d exceeded its particle capacity %u", scale,
                     t->lv[scale].cap);
  return 0;
}

int cuimg_hip_tracker_get_particles(cuimg_hip_tracker* t, int scale, cuimg_particle_f* out, uint32_t cap,
                                    uint32_t* n) {
  CUIMG_TRY(check_scale(t, scale));
  level_counters c;
  CUIMG_TRY(read_counters(t, scale, &c));
  if (n) *n = c.n;
  if (!out) return 0;
  if (c.n > cap) return set_error(t->ctx, CUIMG_E_CAPACITY, "get_particles: %u particles, buffer holds %u", c.n, cap);
  const level_res& R = t->lv[scale];
  CUIMG_HIP(t->ctx, hipMemcpyAsync(out, R.parts[R.cur], sizeof(cuimg_particle_f) * c.n, hipMemcpyDeviceToHost,
                                   t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_set_particles(cuimg_hip_tracker* t, int scale, const cuimg_particle_f* in, uint32_t n) {
  CUIMG_TRY(check_scale(t, scale));
  level_res& R = t->lv[scale];
  if (n > R.cap) return set_error(t->ctx, CUIMG_E_CAPACITY, "set_particles: %u exceeds capacity %u", n, R.cap);
  if (n)
    CUIMG_HIP(t->ctx, hipMemcpyAsync(R.parts[R.cur], in, sizeof(cuimg_particle_f) * n, hipMemcpyHostToDevice,
                                     t->ctx->stream));
  CUIMG_HIP(t->ctx, hipMemcpyAsync(&t->cnt[scale].n, &n, sizeof(uint32_t), hipMemcpyHostToDevice, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_get_matches(cuimg_hip_tracker* t, int scale, uint32_t* out, uint32_t cap, uint32_t* n) {
  CUIMG_TRY(check_scale(t, scale));
  level_counters c;
  CUIMG_TRY(read_counters(t, scale, &c));
  const uint32_t m = c.n_old + c.n_new;
  if (n) *n = m;
  if (!out) return 0;
  if (m > cap) return set_error(t->ctx, CUIMG_E_CAPACITY, "get_matches: %u entries, buffer holds %u", m, cap);
  CUIMG_HIP(t->ctx, hipMemcpyAsync(out, t->lv[scale].matches, sizeof(uint32_t) * m, hipMemcpyDeviceToHost,
                                   t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_compact_has_run(cuimg_hip_tracker* t, int scale, int* out) {
  CUIMG_TRY(check_scale(t, scale));
  *out = t->lv[scale].compact_has_run ? 1 : 0;
  return 0;
}

int cuimg_hip_tracker_sparse(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  CUIMG_TRY(check_scale(t, scale));
  *view = t->lv[scale].sparse;
  return 0;
}

int cuimg_hip_tracker_level(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  CUIMG_TRY(check_scale(t, scale));
  *view = t->lv[scale].u8[(t->frame_cpt + 1) & 1];  // slot of the last frame
  return 0;
}

int cuimg_hip_tracker_saliency(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  CUIMG_TRY(check_scale(t, scale));
  *view = t->lv[scale].sal;
  return 0;
}

int cuimg_hip_tracker_get_decisions(cuimg_hip_tracker* t, int scale, int32_t* accept, float* tr2, float* dist,
                                    uint32_t cap, uint32_t* n) {
  CUIMG_TRY(check_scale(t, scale));
  level_counters c;
  CUIMG_TRY(read_counters(t, scale, &c));
  if (n) *n = c.n_matched;
  if (!accept) return 0;
  if (c.n_matched > cap) return set_error(t->ctx, CUIMG_E_CAPACITY, "get_decisions: buffer too small");
  const level_res& R = t->lv[scale];
  const size_t m = c.n_matched;
  CUIMG_HIP(t->ctx, hipMemcpyAsync(accept, R.dec_accept, 4 * m, hipMemcpyDeviceToHost, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipMemcpyAsync(tr2, R.dec_tr, 8 * m, hipMemcpyDeviceToHost, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipMemcpyAsync(dist, R.dec_dist, 4 * m, hipMemcpyDeviceToHost, t->ctx->stream));
  CUIMG_HIP(t->ctx, hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_matcher_images(cuimg_hip_tracker* t, int scale, cuimg_img* A, cuimg_img* B, cuimg_img* G) {
  CUIMG_TRY(check_scale(t, scale));
  const level_res& R = t->lv[scale];
  const int cur = (t->frame_cpt + 1) & 1, prev = cur ^ 1;
  if (A) *A = t->P.profile == 0 ? R.u8[prev] : R.blur[prev];
  if (B) *B = t->P.profile == 0 ? R.u8[cur] : R.blur[cur];
  if (G) std::memset(G, 0, sizeof(*G));  // computed on the fly, not stored
  return 0;
}

int cuimg_hip_particles_merge(cuimg_hip_ctx* ctx, cuimg_particle_f* parts, uint32_t n, const cuimg_img* sparse) {
  if (!ctx || (!parts && n) || !sparse || sparse->dtype != CUIMG_U32)
    return set_error(ctx, CUIMG_E_INVALID, "particles_merge: bad arguments");
  if (n > kMergeMaxCap)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "particles_merge: %u particles exceed %u", n, kMergeMaxCap);
  const uint32_t cap = std::max<uint32_t>(n, 1);
  void *buf = nullptr, *cnt = nullptr;
  CUIMG_HIP(ctx, hipMalloc(&buf, merge_scratch_bytes(cap)));
  if (hipMalloc(&cnt, sizeof(level_counters)) != hipSuccess) {
    (void)hipFree(buf);
    return set_error(ctx, CUIMG_E_NOMEM, "particles_merge: allocation failed");
  }
  level_counters c{};
  c.n = n;
  int rc = 0;
  if (hipMemcpyAsync(cnt, &c, sizeof c, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = CUIMG_E_NOMEM;
  if (!rc) {
    merge_args m = merge_scratch_carve(buf, cap);
    m.nr = sparse->nrows;
    m.nc = sparse->ncols;
    m.parts = parts;
    m.sparse = static_cast<const uint32_t*>(sparse->begin);
    m.sparse_pitch = sparse->pitch;
    m.cnt = static_cast<level_counters*>(cnt);
    rc = launch_merge_reset(ctx, m);
    if (!rc) rc = launch_merge(ctx, m);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  (void)hipFree(cnt);
  return rc;
}

int cuimg_hip_tracker_work(cuimg_hip_tracker* t, int scale, uint64_t* live, uint64_t* accepted) {
  CUIMG_TRY(check_scale(t, scale));
  level_counters c;
  CUIMG_TRY(read_counters(t, scale, &c));
  if (live) *live = c.live_total;
  if (accepted) *accepted = c.accepted_total;
  return 0;
}

int cuimg_hip_tracker_frame_count(cuimg_hip_tracker* t, int* out) {
  if (!t || !out) return CUIMG_E_INVALID;
  *out = t->frame_cpt;
  return 0;
}

}  // extern "C"
