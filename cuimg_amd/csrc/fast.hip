// FAST-like saliency: standalone launch (the device code and its reference
// notes are in fast_dev.hpp).
#include "fast_dev.hpp"

namespace cuimg {

template <bool CLAMP>
__global__ __launch_bounds__(256) void fast_kernel(const uint8_t* in, int in_pitch, const uint8_t* mask,
                                                   int mask_pitch, uint8_t* out, int out_pitch, int nr, int nc,
                                                   int tiles_x, int n, int th) {
  fast_tile_body<CLAMP>(in, in_pitch, mask, mask_pitch, out, out_pitch, nr, nc, blockIdx.x, tiles_x, n, th);
}

int fast_saliency(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* mask, const cuimg_img* out,
                  int n, int th) {
  if (!in || !out || in->dtype != CUIMG_U8 || out->dtype != CUIMG_U8 ||
      in->nrows != out->nrows || in->ncols != out->ncols || (mask && (mask->dtype != CUIMG_U8 ||
      mask->nrows != in->nrows || mask->ncols != in->ncols)))
    return set_error(ctx, CUIMG_E_INVALID, "fast_saliency: images must be U8 of one domain");
  if (in->nrows <= 0 || in->ncols <= 0) return 0;
  const int tiles_x = ceil_div(in->ncols, kFTW);
  const int blocks = tiles_x * ceil_div(in->nrows, kFTH);
  const uint8_t* m = mask ? static_cast<const uint8_t*>(mask->begin) : nullptr;
  const int mp = mask ? mask->pitch : 0;
  return launch(ctx, "fast_saliency", [&] {
    if (in->border >= 3)
      fast_kernel<false><<<blocks, 256, 0, ctx->stream>>>(
          static_cast<const uint8_t*>(in->begin), in->pitch, m, mp, static_cast<uint8_t*>(out->begin),
          out->pitch, in->nrows, in->ncols, tiles_x, n, th);
    else
      fast_kernel<true><<<blocks, 256, 0, ctx->stream>>>(
          static_cast<const uint8_t*>(in->begin), in->pitch, m, mp, static_cast<uint8_t*>(out->begin),
          out->pitch, in->nrows, in->ncols, tiles_x, n, th);
  });
}

}  // namespace cuimg

using namespace cuimg;

extern "C" int cuimg_hip_fast_saliency(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* mask,
                                       const cuimg_img* out, int n, int th) {
  return fast_saliency(ctx, in, mask, out, n, th);
}
