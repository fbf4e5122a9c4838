// Matrix-core separable Gaussian: kernel instances and launcher (the device
// code and its reference notes are in gaussian_mfma_dev.hpp).
#include "gaussian_mfma_dev.hpp"

namespace cuimg {

template <typename E, int K>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void gauss_mfma_kernel(gauss_jobs jobs, coef_pack cp) {
  int b = xcd_order(blockIdx.x, gridDim.x);
  int k = 0;
  while (k + 1 < jobs.n && b >= jobs.j[k + 1].tile_begin) k++;
  const gauss_job& J = jobs.j[k];
  b -= J.tile_begin;
  __shared__ __align__(16) float lds[gmfma<E, K>::LDS_FLOATS];
  using G = gmfma<E, K>;
  const int tx = b % J.tiles_x, ty = b / J.tiles_x;
  const int s0 = tx * G::SW - G::KA;
  // interior strips: FP32, the staged window inside the image with 16-byte
  // rows; FP32x4, all 16 output pixels inside the image
  const bool interior = G::RGBA ? (tx + 1) * G::SW <= J.nc : J.vec_ok && s0 >= 0 && s0 + G::LWF <= J.nc;
  if (interior && J.stream_out) gauss_mfma_tile<E, K, true, true>(J, tx, ty, cp, lds);
  else if (interior) gauss_mfma_tile<E, K, true, false>(J, tx, ty, cp, lds);
  else gauss_mfma_tile<E, K, false, false>(J, tx, ty, cp, lds);
}

template <typename E>
static void launch_mfma_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp, int K) {
  switch (K) {
#define CUIMG_K(k) \
  case k: gauss_mfma_kernel<E, k><<<blocks, 64, 0, ctx->stream>>>(jobs, cp); break;
    CUIMG_K(1) CUIMG_K(2) CUIMG_K(3) CUIMG_K(4) CUIMG_K(5) CUIMG_K(6) CUIMG_K(7) CUIMG_K(8)
    CUIMG_K(9) CUIMG_K(10) CUIMG_K(11) CUIMG_K(12) CUIMG_K(13) CUIMG_K(14) CUIMG_K(15) CUIMG_K(16)
#undef CUIMG_K
  }
}

// Tiles: 64-column (FP32) or 16-pixel (FP32x4) strips, rows in multiples of
// 16 (th_override rounded up).  2 waves are resident per SIMD (the VGPR budget
// with two blocks in flight per wave).  By default FP32 launches about 2048
// waves, one round; FP32x4, whose strips read 2.5 times their width, about
// 4096, which halves each wave's walk (measured faster on 3840 x 2160).
int convolve_sep_mfma(cuimg_hip_ctx* ctx, gauss_jobs jobs, const coef_pack& cp, int K, bool rgba,
                      int th_override) {
  const int sw = rgba ? 16 : 64;
  long strips = 0;
  for (int i = 0; i < jobs.n; i++) strips += (long)jobs.j[i].nr * ceil_div(jobs.j[i].nc, sw);
  const long want = rgba ? 4096 : 2048;
  int th = th_override > 0 ? th_override : (int)std::min<long>(256, std::max<long>(16, (strips + want - 1) / want));
  th = (th + 15) / 16 * 16;
  int total = 0;
  for (int i = 0; i < jobs.n; i++) {
    gauss_job& J = jobs.j[i];
    J.th = th;
    J.tiles_x = ceil_div(J.nc, sw);
    J.tile_begin = total;
    total += J.tiles_x * ceil_div(J.nr, th);
  }
  return launch(ctx, "gaussian_sep", [&] {
    if (rgba) launch_mfma_k<float4>(ctx, jobs, total, cp, K);
    else launch_mfma_k<float>(ctx, jobs, total, cp, K);
  });
}

}  // namespace cuimg
