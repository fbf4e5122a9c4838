// Context, device memory, image handles, border fill and per-kernel timing.
#include "common.hpp"

#include <cstring>

namespace cuimg {

int set_error(cuimg_hip_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

int hip_error(cuimg_hip_ctx* ctx, hipError_t e, const char* what, const char* file, int line) {
  set_error(ctx, (int)e, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  return (int)e;
}

static hipEvent_t take_event(cuimg_hip_ctx* ctx) {
  if (!ctx->free_events.empty()) {
    hipEvent_t e = ctx->free_events.back();
    ctx->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

launch_scope::launch_scope(cuimg_hip_ctx* c, const char* n) : ctx(c), name(n) {
  if (ctx->profiling) {
    a = take_event(ctx);
    b = take_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
  }
}

int launch_scope::done() {
  hipError_t e = hipGetLastError();
  if (ctx->profiling && a && b) {
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, a, b});
    // harvest periodically so the pending list stays bounded
    if (ctx->pending.size() > 4096) {
      (void)hipStreamSynchronize(ctx->stream);
      for (auto& t : ctx->pending) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        auto& s = ctx->stats[t.name];
        s.first += 1;
        s.second += ms;
        ctx->free_events.push_back(t.a);
        ctx->free_events.push_back(t.b);
      }
      ctx->pending.clear();
    }
  }
  if (e != hipSuccess) return hip_error(ctx, e, name, __FILE__, __LINE__);
  return 0;
}

// ------------------------------------------------------------ border fill
// make_border_clamp_kernel (cuimg/border.h:109-137) restricted to the border
// pixels: every pixel of domain + border outside the domain copies its clamped
// domain pixel.  Threads enumerate the top band, the bottom band, then the
// left and right columns of the domain rows.
template <int ES>
struct elem_t { uint8_t b[ES]; };

template <int ES>
__global__ void fill_border_kernel(char* begin, int pitch, int nr, int nc, int b) {
  const int W = nc + 2 * b;
  const long band = (long)b * W;
  const long total = 2 * band + 2L * b * nr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    int r, c;
    if (i < band) {
      r = -b + (int)(i / W);
      c = -b + (int)(i % W);
    } else if (i < 2 * band) {
      long j = i - band;
      r = nr + (int)(j / W);
      c = -b + (int)(j % W);
    } else {
      long j = i - 2 * band;
      r = (int)(j / (2 * b));
      int k = (int)(j % (2 * b));
      c = k < b ? -b + k : nc + (k - b);
    }
    int rr = clampi(r, 0, nr - 1), cc = clampi(c, 0, nc - 1);
    *reinterpret_cast<elem_t<ES>*>(begin + (ptrdiff_t)r * pitch + (ptrdiff_t)c * ES) =
        *reinterpret_cast<const elem_t<ES>*>(begin + (ptrdiff_t)rr * pitch + (ptrdiff_t)cc * ES);
  }
}

int fill_border(cuimg_hip_ctx* ctx, const cuimg_img* img) {
  if (!img || !img->begin) return set_error(ctx, CUIMG_E_INVALID, "fill_border_clamp: null image");
  const int b = img->border;
  if (b <= 0 || img->nrows <= 0 || img->ncols <= 0) return 0;
  const long total = 2L * b * (img->ncols + 2 * b) + 2L * b * img->nrows;
  int blocks = (int)std::min<long>((total + 255) / 256, 4096);
  char* begin = static_cast<char*>(img->begin);
  return launch(ctx, "fill_border_clamp", [&] {
    switch (dtype_size(img->dtype)) {
      case 1: fill_border_kernel<1><<<blocks, 256, 0, ctx->stream>>>(begin, img->pitch, img->nrows, img->ncols, b); break;
      case 4: fill_border_kernel<4><<<blocks, 256, 0, ctx->stream>>>(begin, img->pitch, img->nrows, img->ncols, b); break;
      case 8: fill_border_kernel<8><<<blocks, 256, 0, ctx->stream>>>(begin, img->pitch, img->nrows, img->ncols, b); break;
      default: fill_border_kernel<16><<<blocks, 256, 0, ctx->stream>>>(begin, img->pitch, img->nrows, img->ncols, b); break;
    }
  });
}

int image_alloc(cuimg_hip_ctx* ctx, int nrows, int ncols, int border, int dtype, cuimg_img* out) {
  const int es = dtype_size(dtype);
  if (!out || es == 0 || nrows < 0 || ncols < 0 || border < 0)
    return set_error(ctx, CUIMG_E_INVALID, "image_alloc: bad arguments (%d x %d, border %d, dtype %d)",
                     nrows, ncols, border, dtype);
  const long lpad = border ? ((long)border * es + 127) / 128 * 128 : 0;
  const long pitch = (lpad + (long)(ncols + border) * es + 127) / 128 * 128;
  const size_t bytes = (size_t)pitch * (size_t)(nrows + 2 * border) + 256;
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, bytes);
  if (e != hipSuccess)
    return set_error(ctx, CUIMG_E_NOMEM, "image_alloc: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  out->begin = static_cast<char*>(base) + (size_t)pitch * border + lpad;
  out->nrows = nrows;
  out->ncols = ncols;
  out->pitch = (int32_t)pitch;
  out->border = border;
  out->dtype = dtype;
  ctx->image_base[out->begin] = base;
  return 0;
}

int image_free(cuimg_hip_ctx* ctx, cuimg_img* img) {
  if (!img || !img->begin) return 0;
  auto it = ctx->image_base.find(img->begin);
  if (it == ctx->image_base.end())
    return set_error(ctx, CUIMG_E_INVALID, "image_free: image not allocated by this context");
  CUIMG_HIP(ctx, hipFree(it->second));
  ctx->image_base.erase(it);
  img->begin = nullptr;
  return 0;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_ctx_create(int device, cuimg_hip_ctx** out) {
  if (!out) return CUIMG_E_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return CUIMG_E_NODEVICE;
  auto* ctx = new cuimg_hip_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return (int)e;
  }
  *out = ctx;
  return 0;
}

int cuimg_hip_ctx_destroy(cuimg_hip_ctx* ctx) {
  if (!ctx) return 0;
  // trackers and batches still alive go first (newest first): they free their
  // own buffers, streams and events through this context
  while (!ctx->children.empty()) {
    const auto child = ctx->children.back();
    ctx->children.pop_back();
    child.second(child.first);
  }
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->image_base) (void)hipFree(kv.second);
  for (auto& t : ctx->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->free_events) (void)hipEventDestroy(e);
  for (auto e : ctx->span)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char* cuimg_hip_last_error(const cuimg_hip_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "no context";
}

int cuimg_hip_sync(cuimg_hip_ctx* ctx) {
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

void* cuimg_hip_stream(cuimg_hip_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cuimg_hip_profile_enable(cuimg_hip_ctx* ctx, int on) {
  ctx->profiling = on != 0;
  return 0;
}

int cuimg_hip_profile_reset(cuimg_hip_ctx* ctx) {
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& t : ctx->pending) {
    ctx->free_events.push_back(t.a);
    ctx->free_events.push_back(t.b);
  }
  ctx->pending.clear();
  ctx->stats.clear();
  return 0;
}

int cuimg_hip_profile_query(cuimg_hip_ctx* ctx, const char* kernel, int64_t* launches,
                            double* total_ms) {
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& t : ctx->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t.a, t.b);
    auto& s = ctx->stats[t.name];
    s.first += 1;
    s.second += ms;
    ctx->free_events.push_back(t.a);
    ctx->free_events.push_back(t.b);
  }
  ctx->pending.clear();
  auto it = ctx->stats.find(kernel ? kernel : "");
  if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.second;
  return 0;
}

int cuimg_hip_timer_start(cuimg_hip_ctx* ctx) {
  if (!ctx) return CUIMG_E_INVALID;
  for (auto& e : ctx->span)
    if (!e) CUIMG_HIP(ctx, hipEventCreate(&e));
  CUIMG_HIP(ctx, hipEventRecord(ctx->span[0], ctx->stream));
  return 0;
}

int cuimg_hip_timer_stop(cuimg_hip_ctx* ctx, double* ms) {
  if (!ctx || !ms || !ctx->span[0]) return CUIMG_E_INVALID;
  CUIMG_HIP(ctx, hipEventRecord(ctx->span[1], ctx->stream));
  CUIMG_HIP(ctx, hipEventSynchronize(ctx->span[1]));
  float t = 0.f;
  CUIMG_HIP(ctx, hipEventElapsedTime(&t, ctx->span[0], ctx->span[1]));
  *ms = t;
  return 0;
}

int cuimg_hip_malloc(cuimg_hip_ctx* ctx, size_t bytes, void** out) {
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess)
    return set_error(ctx, CUIMG_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return 0;
}

int cuimg_hip_free(cuimg_hip_ctx* ctx, void* p) {
  CUIMG_HIP(ctx, hipFree(p));
  return 0;
}

int cuimg_hip_memcpy_h2d(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CUIMG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cuimg_hip_memcpy_d2h(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  CUIMG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cuimg_hip_memcpy_d2d(cuimg_hip_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return CUIMG_E_INVALID;
  return cuimg::launch(ctx, "memcpy_d2d", [&] { (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream); });
}

int cuimg_hip_memset(cuimg_hip_ctx* ctx, void* dst, int value, size_t bytes) {
  CUIMG_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
  return 0;
}

int cuimg_hip_image_alloc(cuimg_hip_ctx* ctx, int nrows, int ncols, int border, int dtype,
                          cuimg_img* out) {
  return image_alloc(ctx, nrows, ncols, border, dtype, out);
}

int cuimg_hip_image_free(cuimg_hip_ctx* ctx, cuimg_img* img) { return image_free(ctx, img); }

int cuimg_hip_image_upload(cuimg_hip_ctx* ctx, const cuimg_img* dst, const void* host,
                           size_t host_pitch) {
  const int es = dtype_size(dst->dtype);
  if (!es || !host) return set_error(ctx, CUIMG_E_INVALID, "image_upload: bad arguments");
  CUIMG_HIP(ctx, hipMemcpy2DAsync(dst->begin, dst->pitch, host, host_pitch, (size_t)dst->ncols * es,
                                  dst->nrows, hipMemcpyHostToDevice, ctx->stream));
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cuimg_hip_image_download(cuimg_hip_ctx* ctx, const cuimg_img* src, void* host,
                             size_t host_pitch) {
  const int es = dtype_size(src->dtype);
  if (!es || !host) return set_error(ctx, CUIMG_E_INVALID, "image_download: bad arguments");
  CUIMG_HIP(ctx, hipMemcpy2DAsync(host, host_pitch, src->begin, src->pitch, (size_t)src->ncols * es,
                                  src->nrows, hipMemcpyDeviceToHost, ctx->stream));
  CUIMG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cuimg_hip_fill_border_clamp(cuimg_hip_ctx* ctx, const cuimg_img* img) {
  return fill_border(ctx, img);
}

}  // extern "C"
