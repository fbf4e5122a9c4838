// Separable Gaussian: host side and standalone launches (the device code and
// its reference notes are in gaussian_dev.hpp, the kernel in
// gaussian_kernel_dev.hpp).
#include "gaussian_dev.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace cuimg {

template <typename TI, typename E, int VEC, int K>
void launch_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp);

template <typename TI, typename E, int VEC>
static int dispatch_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp, int K) {
  switch (K) {
#define CUIMG_K(k) case k: launch_k<TI, E, VEC, k>(ctx, jobs, blocks, cp); return 0;
    CUIMG_K(1) CUIMG_K(2) CUIMG_K(3) CUIMG_K(4) CUIMG_K(5) CUIMG_K(6) CUIMG_K(7) CUIMG_K(8)
    CUIMG_K(9) CUIMG_K(10) CUIMG_K(11) CUIMG_K(12) CUIMG_K(13) CUIMG_K(14) CUIMG_K(15) CUIMG_K(16)
#undef CUIMG_K
  }
  return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported (1..%d)", K, kMaxK);
}

int convolve_sep_mfma(cuimg_hip_ctx* ctx, gauss_jobs jobs, const coef_pack& cp, int K, bool rgba, int th_override);
constexpr int kMfmaMinK = 9;

static int pick_th(int nr, int tiles_x, int total_tiles_other) {
  // aim for >= ~2048 waves in flight over the whole launch, 16..128 rows per tile
  long want = 2048 - total_tiles_other;
  if (want < 256) want = 256;
  long th = ((long)nr * tiles_x + want - 1) / want;
  th = std::max<long>(16, std::min<long>(128, th));
  return (int)th;
}

int convolve_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs, int n,
                       const float* coefs, int K, int th_override) {
  if (n <= 0 || n > kMaxJobs || !ins || !outs || !coefs)
    return set_error(ctx, CUIMG_E_INVALID, "gaussian: bad job list (n=%d)", n);
  if (K < 1 || K > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported (1..%d)", K, kMaxK);
  const int in_dt = ins[0].dtype;
  const int vec = in_dt == CUIMG_F32X4 ? 1 : 4;
  const int W = 64 * vec;
  const int es_in = dtype_size(in_dt), es_out = in_dt == CUIMG_F32X4 ? 16 : 4;
  gauss_jobs jobs{};
  jobs.n = n;
  int total = 0;
  for (int i = 0; i < n; i++) {
    const cuimg_img& a = ins[i];
    const cuimg_img& b = outs[i];
    if (a.dtype != in_dt || !(in_dt == CUIMG_U8 || in_dt == CUIMG_F32 || in_dt == CUIMG_F32X4))
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: input dtype %d not supported", a.dtype);
    const int want_out = in_dt == CUIMG_F32X4 ? CUIMG_F32X4 : CUIMG_F32;
    if (b.dtype != want_out)
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: output dtype %d, expected %d", b.dtype, want_out);
    if (a.nrows != b.nrows || a.ncols != b.ncols || a.nrows <= 0 || a.ncols <= 0 || !a.begin || !b.begin)
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: domains differ or empty (%dx%d vs %dx%d)",
                       a.nrows, a.ncols, b.nrows, b.ncols);
    gauss_job& J = jobs.j[i];
    J.in = a.begin;
    J.out = b.begin;
    J.in_pitch = a.pitch;
    J.out_pitch = b.pitch;
    J.nr = a.nrows;
    J.nc = a.ncols;
    J.tiles_x = ceil_div(a.ncols, W);
    J.vec_ok = ((uintptr_t)a.begin % (vec * es_in) == 0) && (a.pitch % (vec * es_in) == 0) &&
               ((uintptr_t)b.begin % (vec * es_out) == 0) && (b.pitch % (vec * es_out) == 0);
  }
  // FP32 / FP32x4 images with 19 or more taps: the matrix-core kernel (the
  // vector kernel is VALU-bound there; measured faster from K = 9 on)
  if (K >= kMfmaMinK && in_dt != CUIMG_U8) {
    size_t ob = 0;
    for (int i = 0; i < n; i++) ob += (size_t)jobs.j[i].nr * jobs.j[i].nc * es_out;
    for (int i = 0; i < n; i++) jobs.j[i].stream_out = ob > kStreamOutBytes;
    coef_pack cp{};
    for (int m = 0; m < 2 * K + 1; m++) cp.c[m] = coefs[m];
    return convolve_sep_mfma(ctx, jobs, cp, K, in_dt == CUIMG_F32X4, th_override);
  }
  // tile height: sized on the largest image
  int big = 0;
  for (int i = 1; i < n; i++)
    if ((long)jobs.j[i].nr * jobs.j[i].nc > (long)jobs.j[big].nr * jobs.j[big].nc) big = i;
  const int th = th_override > 0 ? th_override : pick_th(jobs.j[big].nr, jobs.j[big].tiles_x, 0);
  for (int i = 0; i < n; i++) {
    gauss_job& J = jobs.j[i];
    J.th = th;
    J.tile_begin = total;
    total += J.tiles_x * ceil_div(J.nr, th);
  }
  // outputs larger than the XCDs' L2s together are streamed past the cache
  size_t out_bytes = 0;
  for (int i = 0; i < n; i++) out_bytes += (size_t)jobs.j[i].nr * jobs.j[i].nc * es_out;
  for (int i = 0; i < n; i++) jobs.j[i].stream_out = out_bytes > kStreamOutBytes;
  coef_pack cp{};
  for (int m = 0; m < 2 * K + 1; m++) cp.c[m] = coefs[m];
  return launch(ctx, "gaussian_sep", [&] {
    int r = 0;
    if (in_dt == CUIMG_U8) r = dispatch_k<uint8_t, float, 4>(ctx, jobs, total, cp, K);
    else if (in_dt == CUIMG_F32) r = dispatch_k<float, float, 4>(ctx, jobs, total, cp, K);
    else r = dispatch_k<float4, float4, 1>(ctx, jobs, total, cp, K);
    (void)r;
  });
}

// gaussian_derivative<0> coefficients as emitted by the reference's generator:
// the double-precision formula rounded to float, printed with 6 significant
// digits and read back as a float literal
// (tools/make_meta_gaussian/make_meta_gaussian.cpp:11-16, 85-89).
int gaussian_coefs(int sigma, int K, float* out) {
  if (sigma < 1 || K < 0 || !out) return CUIMG_E_INVALID;
  const float pi = 3.14159265f;  // CUIMG_PI
  const float s = (float)sigma;
  for (int x = -K; x <= K; x++) {
    double a = 1 / (std::sqrt((double)(2 * pi)) * (double)s);
    float e = (float)(-(x * x)) / (2 * s * s);
    float v = (float)(a * std::exp((double)e));
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", (double)v);
    out[x + K] = std::strtof(buf, nullptr);
  }
  return 0;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_gaussian_coefs(int sigma, int half_width, float* out) {
  return gaussian_coefs(sigma, half_width, out);
}

int cuimg_hip_convolve_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           const float* coefs, int half_width) {
  return convolve_sep_multi(ctx, in, out, 1, coefs, half_width, 0);
}

int cuimg_hip_gaussian_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           int sigma, int half_width) {
  float c[2 * kMaxK + 1];
  if (half_width < 1 || half_width > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported", half_width);
  CUIMG_TRY(gaussian_coefs(sigma, half_width, c));
  return convolve_sep_multi(ctx, in, out, 1, c, half_width, 0);
}

int cuimg_hip_gaussian_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs,
                                 int n, int sigma, int half_width) {
  float c[2 * kMaxK + 1];
  if (half_width < 1 || half_width > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported", half_width);
  CUIMG_TRY(gaussian_coefs(sigma, half_width, c));
  return convolve_sep_multi(ctx, ins, outs, n, c, half_width, 0);
}

// Tile-height override for the autotuner and the roofline sweep.
int cuimg_hip_convolve_sep_th(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                              const float* coefs, int half_width, int th) {
  return convolve_sep_multi(ctx, in, out, 1, coefs, half_width, th);
}

}  // extern "C"
