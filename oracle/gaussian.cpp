// Oracle: Gaussian coefficients, separable convolution and the box pyramid.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace orc {

static const float kPi = 3.14159265f;  // CUIMG_PI, cuimg/util.h:17

// gaussian_derivative<0> as compiled into the CPU overloads
// (cuimg/gaussian_kernel.h:13-18): every operation in float.
float gauss_coef_runtime(float sigma, int x) {
  float num = (float)(-(x * x));
  return (1.f / (sqrtf(2.f * kPi) * sigma)) * expf(num / (2.f * sigma * sigma));
}

// The coefficient the GPU meta-template kernels use: the generator
// (tools/make_meta_gaussian/make_meta_gaussian.cpp:11-16) calls the unqualified
// ::sqrt / ::exp, i.e. the double-precision C functions, on float arguments, and
// returns the product as float; the value is then printed with the default
// ostream precision of 6 significant digits (:85-89) and compiled back as a
// float literal.  Sigma is an integer template parameter (:63,168).
float gauss_coef_meta(int sigma_i, int x) {
  float sigma = (float)sigma_i;
  float two_pi = 2 * kPi;
  double a = 1 / (std::sqrt((double)two_pi) * (double)sigma);
  float expo = (float)(-(x * x)) / (2 * sigma * sigma);
  float res = (float)(a * std::exp((double)expo));
  char buf[64];
  std::snprintf(buf, sizeof buf, "%g", (double)res);
  return std::strtof(buf, nullptr);
}

void gauss_coefs(int sigma, int K, bool meta_rounding, float* out) {
  for (int x = -K; x <= K; x++)
    out[x + K] = meta_rounding ? gauss_coef_meta(sigma, x) : gauss_coef_runtime((float)sigma, x);
}

// Separable convolution with the semantics of the GPU meta kernels
// (cuimg/gpu/meta_convolve.h:25-128): row pass then column pass, taps
// x = -K..K, clamp-to-edge reads (texture clamp addressing), FP32 intermediate.
// The recursive template  v_R*c_R + iter<R+1>  associates right to left:
//   c_{-K}v_{-K} + (c_{-K+1}v_{-K+1} + (... + c_K v_K)).
// fused: every term after the first (tap +K) is added with one fused
// multiply-add, s = fma(c_j, v_j, s) (the GPU's arithmetic); otherwise each
// product is rounded before it is added.
template <typename TI>
void gaussian_sep_f32(const image<TI>& in, const float* coef, int K, image<float>& out, bool fused) {
  const int nr = in.nr, nc = in.nc;
  image<float> tmp(nr, nc);
#pragma omp parallel for schedule(static, 2)
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      float s = (float)in.clamped(r, c + K) * coef[2 * K];
      for (int j = K - 1; j >= -K; j--)
        s = fused ? std::fmaf(coef[j + K], (float)in.clamped(r, c + j), s) : (float)in.clamped(r, c + j) * coef[j + K] + s;
      tmp(r, c) = s;
    }
  out = image<float>(nr, nc);
#pragma omp parallel for schedule(static, 2)
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      float s = tmp.clamped(r + K, c) * coef[2 * K];
      for (int j = K - 1; j >= -K; j--)
        s = fused ? std::fmaf(coef[j + K], tmp.clamped(r + j, c), s) : tmp.clamped(r + j, c) * coef[j + K] + s;
      out(r, c) = s;
    }
}
template void gaussian_sep_f32<float>(const image<float>&, const float*, int, image<float>&, bool);
template void gaussian_sep_f32<uint8_t>(const image<uint8_t>&, const float*, int, image<float>&, bool);

// FP64 twin of the same operation order, used to bound FP32 error.
void gaussian_sep_f64(const image<double>& in, const double* coef, int K, image<double>& out) {
  const int nr = in.nr, nc = in.nc;
  image<double> tmp(nr, nc);
#pragma omp parallel for schedule(static, 2)
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      double s = in.clamped(r, c + K) * coef[2 * K];
      for (int j = K - 1; j >= -K; j--) s = in.clamped(r, c + j) * coef[j + K] + s;
      tmp(r, c) = s;
    }
  out = image<double>(nr, nc);
#pragma omp parallel for schedule(static, 2)
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      double s = tmp.clamped(r + K, c) * coef[2 * K];
      for (int j = K - 1; j >= -K; j--) s = tmp.clamped(r + j, c) * coef[j + K] + s;
      out(r, c) = s;
    }
}

// mipmap_kernel (cuimg/gpu/mipmap.h:25-50) driven by subsample (:206-227):
// out(p) = (in(2p) + in(2p+(0,1)) + in(2p+(1,0)) + in(2p+(1,1))) / 4.
// gl8u + gl8u promotes to int (cuimg/meta.h:85-118), so the u8 mean is an
// integer division; the output domain is obox2d / 2 = floor (cuimg/obox2d.h:69-72).
void pyr_down_u8(const image<uint8_t>& in, image<uint8_t>& out) {
  out = image<uint8_t>(in.nr / 2, in.nc / 2);
  for (int r = 0; r < out.nr; r++)
    for (int c = 0; c < out.nc; c++) {
      int s = int(in(2 * r, 2 * c)) + int(in(2 * r, 2 * c + 1)) + int(in(2 * r + 1, 2 * c)) +
              int(in(2 * r + 1, 2 * c + 1));
      out(r, c) = (uint8_t)(s / 4);
    }
}

// Same kernel on float images: ((a + b) + c) + d, then / 4.
void pyr_down_f32(const image<float>& in, image<float>& out) {
  out = image<float>(in.nr / 2, in.nc / 2);
  for (int r = 0; r < out.nr; r++)
    for (int c = 0; c < out.nc; c++) {
      float s = in(2 * r, 2 * c) + in(2 * r, 2 * c + 1);
      s = s + in(2 * r + 1, 2 * c);
      s = s + in(2 * r + 1, 2 * c + 1);
      out(r, c) = s / 4;
    }
}

}  // namespace orc
