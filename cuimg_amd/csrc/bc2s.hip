// Standalone entry points of the bc2s path: the feature images, the integer
// Gaussians, the descriptor and the gradient-descent matcher on device arrays.
#include "bc2s_dev.hpp"

namespace cuimg {

__global__ __launch_bounds__(256) void bc2s_prep_kernel(bc2s_prep_args a) { bc2s_prep_tile(a, blockIdx.x); }

__global__ __launch_bounds__(256) void gint_kernel(gint_args a, int columns) {
  const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
  if (c < a.nc) gint_pass(a, columns != 0, r, c);
}

__global__ __launch_bounds__(256) void bc2s_features_kernel(bc2s_view F, const int2* pts, uint32_t n, uint4* out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = bc2s_feature_at(F, pts[i].x, pts[i].y);
}

__global__ __launch_bounds__(256) void bc2s_match_kernel(bc2s_view F, const uint4* feats, const int2* preds,
                                                         uint32_t n, int scale, int2* disp, float* dist) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const descent_out m = scale == 0 ? bc2s_two_step(F, feats[i], preds[i].x, preds[i].y)
                                   : bc2s_descent(F, feats[i], preds[i].x, preds[i].y, scale == 1);
  disp[i] = make_int2(m.dr, m.dc);
  dist[i] = m.dist;
}

static bool is_u8(const cuimg_img* a) { return a && a->begin && a->dtype == CUIMG_U8 && a->nrows > 0 && a->ncols > 0; }
static bool same_shape(const cuimg_img* a, const cuimg_img* b) { return a->nrows == b->nrows && a->ncols == b->ncols; }

static int check_feature_images(cuimg_hip_ctx* ctx, const cuimg_img* s1, const cuimg_img* s2, const char* who,
                                bc2s_view* F) {
  if (!is_u8(s1) || !is_u8(s2) || !same_shape(s1, s2) || s1->border < 3 || s2->border < 6)
    return set_error(ctx, CUIMG_E_INVALID, "%s: s1 / s2 must be U8 images of one shape with clamp-filled borders "
                     ">= 3 / >= 6", who);
  *F = bc2s_view{static_cast<const uint8_t*>(s1->begin), static_cast<const uint8_t*>(s2->begin), s1->pitch,
                 s2->pitch, s1->nrows, s1->ncols};
  return 0;
}

static int prep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* s1, const cuimg_img* s2,
                const cuimg_img* contrast, const char* who) {
  if (!ctx) return CUIMG_E_INVALID;
  if (!is_u8(in) || !is_u8(s1) || !same_shape(in, s1) || (s2 && (!is_u8(s2) || !same_shape(in, s2))) ||
      s1->border > 3 || (s2 && s2->border > 6) ||
      (contrast && (!contrast->begin || !same_shape(in, contrast) ||
                    (contrast->dtype != CUIMG_U8 && contrast->dtype != CUIMG_U32))))
    return set_error(ctx, CUIMG_E_INVALID, "%s: U8 input, U8 outputs of its shape (borders <= 3 / 6), contrast "
                     "U8 / U32 or NULL", who);
  if (s1->begin == in->begin)
    return set_error(ctx, CUIMG_E_INVALID, "%s: the output may not alias the input", who);
  bc2s_prep_args a{};
  a.in = static_cast<const uint8_t*>(in->begin);
  a.in_pitch = in->pitch;
  if (contrast && contrast->dtype == CUIMG_U8) {
    a.con = static_cast<uint8_t*>(contrast->begin);
    a.con_pitch = contrast->pitch;
  }
  if (contrast && contrast->dtype == CUIMG_U32) {
    a.con32 = static_cast<uint32_t*>(contrast->begin);
    a.con32_pitch = contrast->pitch;
  }
  a.s1 = static_cast<uint8_t*>(s1->begin);
  a.s1_pitch = s1->pitch;
  a.s1_border = s1->border;
  if (s2) {
    a.s2 = static_cast<uint8_t*>(s2->begin);
    a.s2_pitch = s2->pitch;
    a.s2_border = s2->border;
  }
  a.nr = in->nrows;
  a.nc = in->ncols;
  a.tiles_x = ceil_div(a.nc, kB2W);
  const int tiles = a.tiles_x * ceil_div(a.nr, kB2H);
  return launch(ctx, "bc2s_images", [&] { bc2s_prep_kernel<<<tiles, 256, 0, ctx->stream>>>(a); });
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_bc2s_images(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* s1, const cuimg_img* s2,
                          const cuimg_img* contrast) {
  if (!s2) return set_error(ctx, CUIMG_E_INVALID, "bc2s_images: s2 is required");
  return prep(ctx, in, s1, s2, contrast, "bc2s_images");
}

int cuimg_hip_blur121(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out) {
  return prep(ctx, in, out, nullptr, nullptr, "blur121");
}

int cuimg_hip_gaussian_int(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out, float sigma, int hw) {
  if (!ctx) return CUIMG_E_INVALID;
  if (!is_u8(in) || !is_u8(out) || !same_shape(in, out) || in->begin == out->begin)
    return set_error(ctx, CUIMG_E_INVALID, "gaussian_int: U8 input and a separate U8 output of its shape");
  if (hw < 0 || hw > kGiMaxHw || !(sigma > 0.f))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian_int: half width %d not in 0..%d or sigma <= 0", hw,
                     kGiMaxHw);
  gint_args a{};
  int sum = 0;
  for (int i = -hw; i <= hw; i++) {
    // gaussian_derivative<0>(sigma, i) * 100 truncated (gpu/gaussian_blur.h:41-44)
    const float g = (1.f / (sqrtf(2.f * 3.14159265f) * sigma)) * expf((float)(-(i * i)) / (2.f * sigma * sigma));
    a.w[i + hw] = (int)(g * 100.f);
    sum += a.w[i + hw];
  }
  if (sum <= 0) return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian_int: the integer taps sum to 0");
  cuimg_img tmp{};
  CUIMG_TRY(image_alloc(ctx, in->nrows, in->ncols, 0, CUIMG_U8, &tmp));
  a.in = static_cast<const uint8_t*>(in->begin);
  a.in_pitch = in->pitch;
  a.tmp = static_cast<uint8_t*>(tmp.begin);
  a.tmp_pitch = tmp.pitch;
  a.out = static_cast<uint8_t*>(out->begin);
  a.out_pitch = out->pitch;
  a.nr = in->nrows;
  a.nc = in->ncols;
  a.hw = hw;
  a.sum = sum;
  const dim3 grid(ceil_div(a.nc, 256), a.nr);
  int rc = launch(ctx, "gaussian_int", [&] {
    gint_kernel<<<grid, 256, 0, ctx->stream>>>(a, 1);
    gint_kernel<<<grid, 256, 0, ctx->stream>>>(a, 0);
  });
  if (!rc) {
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_error(ctx, e, "gaussian_int", __FILE__, __LINE__);
  }
  image_free(ctx, &tmp);
  return rc;
}

int cuimg_hip_bc2s_features(cuimg_hip_ctx* ctx, const cuimg_img* s1, const cuimg_img* s2, const int32_t* pts,
                            uint32_t n, uint8_t* out16) {
  if (!ctx || (n && (!pts || !out16))) return CUIMG_E_INVALID;
  bc2s_view F;
  CUIMG_TRY(check_feature_images(ctx, s1, s2, "bc2s_features", &F));
  if (!n) return 0;
  return launch(ctx, "bc2s_features", [&] {
    bc2s_features_kernel<<<ceil_div((int)n, 256), 256, 0, ctx->stream>>>(F, reinterpret_cast<const int2*>(pts), n,
                                                                         reinterpret_cast<uint4*>(out16));
  });
}

int cuimg_hip_bc2s_match(cuimg_hip_ctx* ctx, const cuimg_img* s1, const cuimg_img* s2, const uint8_t* feats16,
                         const int32_t* preds, uint32_t n, int scale, int32_t* out_disp, float* out_dist) {
  if (!ctx || (n && (!feats16 || !preds || !out_disp || !out_dist)) || scale < 0 || scale > 2) return CUIMG_E_INVALID;
  bc2s_view F;
  CUIMG_TRY(check_feature_images(ctx, s1, s2, "bc2s_match", &F));
  if (!n) return 0;
  return launch(ctx, "bc2s_match", [&] {
    bc2s_match_kernel<<<ceil_div((int)n, 256), 256, 0, ctx->stream>>>(
        F, reinterpret_cast<const uint4*>(feats16), reinterpret_cast<const int2*>(preds), n, scale,
        reinterpret_cast<int2*>(out_disp), out_dist);
  });
}

}  // extern "C"
