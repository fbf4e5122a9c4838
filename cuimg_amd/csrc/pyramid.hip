// Box pyramid: standalone launch (the device code and its reference notes are
// in pyramid_dev.hpp).
#include "pyramid_dev.hpp"

#include <algorithm>

namespace cuimg {

template <typename T>
__global__ __launch_bounds__(kPyrThreads) void pyramid_kernel(pyr_args a) {
  pyramid_tile<T>(a, blockIdx.x);
}

// Workgroups wanted per launch (about 4 per CU) before tiles get narrower.
constexpr int kPyrMinTiles = 960;

int pyr_tiling(pyr_args& a, int nr, int nc, int es) {
  int tile_log2 = 0;
  while ((1 << (tile_log2 + 1)) <= kPyrTileBytes / es) tile_log2++;  // elements per tile
  const int thl = std::max(4, a.n - 1);          // >= 16 rows and 2^(n-1)
  int twl = tile_log2 - thl;                     // widest 16 KiB tile
  auto tiles = [&](int w) { return (long)ceil_div(nc, 1 << w) * ceil_div(nr, 1 << thl); };
  while (twl > 6 && twl > a.n - 1 && tiles(twl) < kPyrMinTiles) twl--;
  a.tw_log2 = twl;
  a.th_log2 = thl;
  a.tiles_x = ceil_div(nc, 1 << twl);
  return (int)tiles(twl);
}

static bool aligned16(const void* p, int pitch) {
  return ((uintptr_t)p % 16 == 0) && (pitch % 16 == 0);
}

int pyramid_build(cuimg_hip_ctx* ctx, const cuimg_img* frame, const cuimg_img* levels, int n) {
  if (!frame || !levels || n < 1 || n > 7)
    return set_error(ctx, CUIMG_E_INVALID, "pyramid_build: bad arguments (n=%d, at most 7 levels)", n);
  const int dt = levels[0].dtype;
  if (dt != CUIMG_U8 && dt != CUIMG_F32)
    return set_error(ctx, CUIMG_E_INVALID, "pyramid_build: dtype %d not supported", dt);
  if (frame->dtype != dt || frame->nrows != levels[0].nrows || frame->ncols != levels[0].ncols)
    return set_error(ctx, CUIMG_E_INVALID, "pyramid_build: frame does not match level 0");
  pyr_args a{};
  a.frame = frame->begin;
  a.frame_pitch = frame->pitch;
  a.copy_level0 = frame->begin != levels[0].begin;
  a.n = n;
  for (int l = 0; l < n; l++) {
    const cuimg_img& L = levels[l];
    if (L.dtype != dt || !L.begin)
      return set_error(ctx, CUIMG_E_INVALID, "pyramid_build: level %d has dtype %d", l, L.dtype);
    if (l > 0 && (L.nrows != levels[l - 1].nrows / 2 || L.ncols != levels[l - 1].ncols / 2))
      return set_error(ctx, CUIMG_E_INVALID, "pyramid_build: level %d is %dx%d, expected %dx%d", l,
                       L.nrows, L.ncols, levels[l - 1].nrows / 2, levels[l - 1].ncols / 2);
    a.lv[l] = pyr_level{L.begin, L.pitch, L.nrows, L.ncols, L.border};
  }
  a.frame_vec_ok = aligned16(frame->begin, frame->pitch);
  a.l0_vec_ok = aligned16(levels[0].begin, levels[0].pitch);
  const int es = dt == CUIMG_U8 ? 1 : 4;
  a.stream_out = (size_t)levels[0].nrows * levels[0].ncols * es > kStreamOutBytes;
  const int blocks = pyr_tiling(a, levels[0].nrows, levels[0].ncols, es);
  if (blocks == 0) return 0;
  return launch(ctx, "pyramid_build", [&] {
    if (dt == CUIMG_U8) pyramid_kernel<uint8_t><<<blocks, kPyrThreads, 0, ctx->stream>>>(a);
    else pyramid_kernel<float><<<blocks, kPyrThreads, 0, ctx->stream>>>(a);
  });
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_pyramid_build(cuimg_hip_ctx* ctx, const cuimg_img* frame, const cuimg_img* levels,
                            int nlevels) {
  return pyramid_build(ctx, frame, levels, nlevels);
}

int cuimg_hip_pyr_down(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out) {
  // a two-level pyramid whose level 0 is the input itself
  // subsample writes the output domain only: neither border is touched
  cuimg_img lv[2] = {*in, *out};
  lv[0].border = 0;
  lv[1].border = 0;
  return pyramid_build(ctx, in, lv, 2);
}

}  // extern "C"
