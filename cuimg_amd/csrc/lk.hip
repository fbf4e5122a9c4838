// Pyramid LK matcher: the standalone launch of one scale's particles
// (pyrlk_match_particles_kernel, tracking2/pyrlk_tracker.hpp:676-769).  The
// device code and its reference notes are in lk_dev.hpp; the tracker's frame
// step runs the same body as one job of a batched launch.
#include "lk_dev.hpp"

#include <algorithm>
#include <mutex>

namespace cuimg {

template <typename TI, int NPL, bool SEL>
__global__ __launch_bounds__(kLkThreads) __attribute__((amdgpu_waves_per_eu(4))) void lk_kernel(lk_args a, int ppw) {
  lk_body<TI, NPL, SEL>(a, blockIdx.x, gridDim.x, ppw);
}

// The matcher's counters: all zero but the particle count (a fill kernel, no
// host copy, so the allocation-free entry point needs no staging either).
__global__ void lk_counters_init_kernel(level_counters* c, uint32_t n) {
  if (threadIdx.x == 0) {
    level_counters z{};
    z.n = n;
    *c = z;
  }
}

// Device scratch of cuimg_hip_lk_track_noalloc: the counters block, then the
// decision buffers (tr2, accept, dist) of up to kNoAllocMaxParticles particles
// when the caller passes none.  One copy per GPU, held until the process ends.
constexpr uint32_t kNoAllocMaxParticles = 1u << 18;
constexpr size_t kNoAllocCounterBytes = 256;
constexpr size_t kNoAllocScratchBytes = kNoAllocCounterBytes + (size_t)kNoAllocMaxParticles * 16;
static_assert(sizeof(level_counters) <= kNoAllocCounterBytes, "counters block");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the context's device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept.
constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void* scratch_ptr[kMaxDevices];

static int scratch_address(cuimg_hip_ctx* ctx, void** out) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk_track_noalloc: device %d", dev);
  if (!scratch_ptr[dev]) {
    int cur = 0;
    CUIMG_HIP(ctx, hipGetDevice(&cur));
    if (cur != dev) CUIMG_HIP(ctx, hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) CUIMG_HIP(ctx, hipSetDevice(cur));
    CUIMG_HIP(ctx, e);
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return 0;
}

// Shared by both entry points.  scratch_dev: the device scratch to use, or
// NULL to allocate (and free) it for this call.
static int lk_track_impl(cuimg_hip_ctx* ctx, const cuimg_img* prev, const cuimg_img* cur, int nlevels, int scale,
                         cuimg_particle_f* parts, uint32_t n, const cuimg_lk_params* p, const cuimg_img* sparse,
                         int32_t* accept, float* tr2, float* dist, void* scratch_dev) {
  if (!ctx || !prev || !cur || !p || (n && !parts)) return CUIMG_E_INVALID;
  if (nlevels < 1 || nlevels > kMaxLevels || scale < 0 || scale >= nlevels)
    return set_error(ctx, CUIMG_E_INVALID, "lk_track: %d levels, scale %d", nlevels, scale);
  if (p->ws < 1 || p->ws > 15 || !(p->ws & 1) || p->max_iters < 0 || p->top_n < 1)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk_track: window %d must be odd and <= 15", p->ws);
  if (accept && (!tr2 || !dist)) return set_error(ctx, CUIMG_E_INVALID, "lk_track: accept, tr2 and dist go together");
  const int dt = prev[0].dtype;
  if (dt != CUIMG_U8 && dt != CUIMG_F32) return set_error(ctx, CUIMG_E_INVALID, "lk_track: images must be U8 or F32");
  lk_args a{};
  for (int j = 0; j < nlevels; j++) {
    const cuimg_img& A = prev[j];
    const cuimg_img& B = cur[j];
    if (A.dtype != dt || B.dtype != dt || A.nrows != B.nrows || A.ncols != B.ncols || !A.begin || !B.begin ||
        A.nrows < 1 || A.ncols < 1)
      return set_error(ctx, CUIMG_E_INVALID, "lk_track: level %d images differ or are empty", j);
    // patches are addressed with 32-bit byte offsets from the image start
    if (A.pitch < 1 || B.pitch < 1 || A.pitch >= (1 << 24) || B.pitch >= (1 << 24) || A.nrows >= (1 << 24) ||
        (uint64_t)A.nrows * (uint64_t)std::max(A.pitch, B.pitch) >= (1ull << 32))
      return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk_track: level %d too large for 32-bit offsets", j);
    a.lv[j] = lk_level_view{A.begin, B.begin, A.pitch, B.pitch, A.nrows, A.ncols};
  }
  const cuimg_img& S = prev[scale];
  if (sparse && (sparse->dtype != CUIMG_U32 || sparse->nrows != S.nrows || sparse->ncols != S.ncols))
    return set_error(ctx, CUIMG_E_INVALID, "lk_track: sparse image must be U32 of the scale's domain");
  if (scratch_dev && !accept && n > kNoAllocMaxParticles)
    return set_error(ctx, CUIMG_E_CAPACITY,
                     "lk_track_noalloc: %u particles; without decision buffers at most %u (the fixed scratch)", n,
                     kNoAllocMaxParticles);
  if (n == 0) return 0;
  a.nlevels = nlevels;
  a.s = scale;
  a.parts = parts;
  a.sparse = sparse ? static_cast<uint32_t*>(sparse->begin) : nullptr;
  a.sparse_pitch = sparse ? sparse->pitch : 0;
  a.pd_r1 = 3;
  a.pd_c1 = 3;
  a.pd_r2 = S.nrows - 4;
  a.pd_c2 = S.ncols - 4;
  a.P = make_lk_params(p->ws, p->max_iters, std::min(p->top_n, p->ws * p->ws), p->eps, p->k, p->template_u8);
  // counters and, when not given, decision buffers
  void* tmp = scratch_dev;
  if (!tmp) {
    const size_t need = kNoAllocCounterBytes + (accept ? 0 : (size_t)n * 16);
    CUIMG_HIP(ctx, hipMalloc(&tmp, need));
  }
  a.cnt = static_cast<level_counters*>(tmp);
  char* q = static_cast<char*>(tmp) + kNoAllocCounterBytes;
  a.dec_tr = accept ? reinterpret_cast<float2*>(tr2) : reinterpret_cast<float2*>(q);
  a.dec_accept = accept ? accept : reinterpret_cast<int32_t*>(q + 8 * (size_t)n);
  a.dec_dist = accept ? dist : reinterpret_cast<float*>(q + 12 * (size_t)n);
  int rc = launch(ctx, "lk_counters_init", [&] { lk_counters_init_kernel<<<1, 64, 0, ctx->stream>>>(a.cnt, n); });
  if (!rc) {
    const int ppw = 8;
    const int blocks = (int)((n + kLkWaves * ppw - 1) / (kLkWaves * ppw));
    rc = launch(ctx, "lk_track", [&] {
      lk_variant(a.P.ws, a.P.top_n, dt == CUIMG_U8, [&](auto ti, auto npl, auto sel) {
        using TI = typename decltype(ti)::type;
        lk_kernel<TI, decltype(npl)::value, decltype(sel)::value><<<blocks, kLkThreads, 0, ctx->stream>>>(a, ppw);
      });
    });
  }
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (!scratch_dev) (void)hipFree(tmp);
  if (!rc && es != hipSuccess) rc = hip_error(ctx, es, "hipStreamSynchronize", __FILE__, __LINE__);
  return rc;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" int cuimg_hip_lk_track(cuimg_hip_ctx* ctx, const cuimg_img* prev, const cuimg_img* cur, int nlevels,
                                  int scale, cuimg_particle_f* parts, uint32_t n, const cuimg_lk_params* p,
                                  const cuimg_img* sparse, int32_t* accept, float* tr2, float* dist) {
  return lk_track_impl(ctx, prev, cur, nlevels, scale, parts, n, p, sparse, accept, tr2, dist, nullptr);
}

extern "C" int cuimg_hip_lk_track_noalloc(cuimg_hip_ctx* ctx, const cuimg_img* prev, const cuimg_img* cur,
                                          int nlevels, int scale, cuimg_particle_f* parts, uint32_t n,
                                          const cuimg_lk_params* p, const cuimg_img* sparse, int32_t* accept,
                                          float* tr2, float* dist) {
  if (!ctx) return CUIMG_E_INVALID;
  void* dev = nullptr;
  CUIMG_TRY(scratch_address(ctx, &dev));
  return lk_track_impl(ctx, prev, cur, nlevels, scale, parts, n, p, sparse, accept, tr2, dist, dev);
}
