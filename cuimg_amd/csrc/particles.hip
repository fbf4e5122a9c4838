// Particle container for gfx950: standalone launches of the container and
// merge kernels (the batched frame step lives in tracker.hip).
#include "particles_dev.hpp"

namespace cuimg {

__global__ void fill_u32_kernel(uint32_t* begin, int pitch, int r0, int r1, int c0, int c1, uint32_t v) {
  fill_u32_body(begin, pitch, r0, r1, c0, c1, v, blockIdx.x, gridDim.x);
}

int launch_fill_u32(cuimg_hip_ctx* ctx, uint32_t* begin, int pitch, int r0, int r1, int c0, int c1, uint32_t v) {
  const long total = (long)(r1 - r0) * (c1 - c0);
  if (total <= 0) return 0;
  // one workgroup per row, or (a single row) one per 1024 16-byte stores
  const int blocks = r1 - r0 > 1 ? std::min(r1 - r0, 8192) : (int)std::min<long>((total + 4095) / 4096, 8192);
  return launch(ctx, "fill_u32", [&] {
    fill_u32_kernel<<<blocks, 256, 0, ctx->stream>>>(begin, pitch, r0, r1, c0, c1, v);
  });
}

size_t merge_scratch_bytes(uint32_t cap) {
  return (size_t)pow2_at_least(cap) * 8 + (size_t)cap * sizeof(merge_actor) + 2 * (size_t)cap * 4 + 256;
}

merge_args merge_scratch_carve(void* base, uint32_t cap) {
  const uint32_t pc = pow2_at_least(cap);
  merge_args m{};
  char* p = static_cast<char*>(base);
  m.pend = reinterpret_cast<unsigned long long*>(p);
  p += (size_t)pc * 8;
  m.actors = reinterpret_cast<merge_actor*>(p);
  p += (size_t)cap * sizeof(merge_actor);
  m.zeroer = reinterpret_cast<uint32_t*>(p);
  p += (size_t)cap * 4;
  m.minunres = reinterpret_cast<uint32_t*>(p);
  m.cap = cap;
  m.pend_cap = pc;
  return m;
}

int launch_merge_reset(cuimg_hip_ctx* ctx, const merge_args& a) {
  // zeroer and minunres are adjacent: one fill of 2*cap words
  return launch_fill_u32(ctx, a.zeroer, 0, 0, 1, 0, (int)(2 * a.cap), kNoIdx);
}

__global__ __launch_bounds__(256) void merge_scan_kernel(merge_args a) { merge_scan_body(a, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(256) void merge_round1_kernel(merge_args a) {
  merge_round1_body(a, blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(1024) void merge_resolve_kernel(merge_args a) { merge_resolve_body(a); }

int launch_merge(cuimg_hip_ctx* ctx, const merge_args& a) {
  const int blocks = (int)std::min<uint32_t>((a.cap + 255) / 256, 2048);
  CUIMG_TRY(launch(ctx, "merge_scan", [&] { merge_scan_kernel<<<blocks, 256, 0, ctx->stream>>>(a); }));
  CUIMG_TRY(launch(ctx, "merge_round1", [&] { merge_round1_kernel<<<blocks, 256, 0, ctx->stream>>>(a); }));
  return launch(ctx, "merge_resolve", [&] { merge_resolve_kernel<<<1, 1024, 0, ctx->stream>>>(a); });
}

}  // namespace cuimg
