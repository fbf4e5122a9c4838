// The standalone separable-Gaussian kernel and its launcher.  The kernel
// instances (16 half widths x 3 pixel types) are compiled in four units,
// gaussian_k0.hip .. gaussian_k3.hip, one per K mod 4; gaussian.hip holds the
// host side.
#pragma once

#include "gaussian_dev.hpp"

#include <type_traits>

namespace cuimg {

// Rows in flight per wave on the direct-to-LDS path.  Up to 13 taps three
// rows suffice (2 waves per SIMD at the default tile heights); deeper rings
// measured 1-3 % slower, 12 and more rows much slower on FP32x4 (LDS).
template <int K>
constexpr int dma_depth() {
  return K <= 6 ? 3 : 6;
}

template <typename TI, typename E, int VEC, int K>
__global__ __launch_bounds__(64) void gauss_sep_kernel(gauss_jobs jobs, coef_pack cp) {
  int b = xcd_order(blockIdx.x, gridDim.x);
  int k = 0;
  while (k + 1 < jobs.n && b >= jobs.j[k + 1].tile_begin) k++;
  const gauss_job& J = jobs.j[k];
  b -= J.tile_begin;
  const int tx = b % J.tiles_x, ty = b / J.tiles_x;
  // one LDS buffer for whichever tile code runs
  constexpr int kDmaBytes = dma_ring<E, VEC, K, dma_depth<K>()>::bytes;
  constexpr int kTileBytes = gauss_tile_lds_bytes<E, VEC, K>();
  constexpr bool kDma = std::is_same<TI, E>::value && sizeof(E) * VEC == 16;
  __shared__ __align__(16) char lds_mem[kDma && kDmaBytes > kTileBytes ? kDmaBytes : kTileBytes];
  if constexpr (kDma) {
    if (J.vec_ok && (tx + 1) * 64 * VEC <= J.nc) {
      if (ty & 1) gauss_tile_dma<E, VEC, K, dma_depth<K>(), true>(J, tx, ty, cp, lds_mem);
      else gauss_tile_dma<E, VEC, K, dma_depth<K>(), false>(J, tx, ty, cp, lds_mem);
      return;
    }
  }
  gauss_tile<TI, E, VEC, K>(J, tx, ty, cp, lds_mem);
}

// ------------------------------------------------------------------ host
template <typename TI, typename E, int VEC, int K>
void launch_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp) {
  gauss_sep_kernel<TI, E, VEC, K><<<blocks, 64, 0, ctx->stream>>>(jobs, cp);
}

// Explicit instances of launch_k for one half width, all pixel types.
#define CUIMG_GAUSS_INSTANCES(k)                                                                       \
  template void launch_k<uint8_t, float, 4, k>(cuimg_hip_ctx*, const gauss_jobs&, int, const coef_pack&); \
  template void launch_k<float, float, 4, k>(cuimg_hip_ctx*, const gauss_jobs&, int, const coef_pack&);   \
  template void launch_k<float4, float4, 1, k>(cuimg_hip_ctx*, const gauss_jobs&, int, const coef_pack&);

}  // namespace cuimg
