// Shared host/device helpers of the cuimg HIP library (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/cuimg_hip.h"

struct cuimg_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  // allocation bookkeeping: image begin -> hipMalloc base
  std::map<void*, void*> image_base;
  // per-kernel event timing
  bool profiling = false;
  struct timed_launch {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<timed_launch> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<int64_t, double>> stats;
  // span timer (cuimg_hip_timer_start / _stop)
  hipEvent_t span[2] = {nullptr, nullptr};
  // trackers and batches made from this context and not yet destroyed, with
  // their destroy functions: cuimg_hip_ctx_destroy destroys them first
  std::vector<std::pair<void*, void (*)(void*)>> children;
};

namespace cuimg {

int set_error(cuimg_hip_ctx* ctx, int code, const char* fmt, ...);
int hip_error(cuimg_hip_ctx* ctx, hipError_t e, const char* what, const char* file, int line);

#define CUIMG_HIP(ctx, expr)                                                   \
  do {                                                                         \
    hipError_t e__ = (expr);                                                   \
    if (e__ != hipSuccess) return ::cuimg::hip_error(ctx, e__, #expr, __FILE__, __LINE__); \
  } while (0)

#define CUIMG_TRY(expr)          \
  do {                           \
    int r__ = (expr);            \
    if (r__ != 0) return r__;    \
  } while (0)

inline int dtype_size(int dtype) {
  switch (dtype) {
    case CUIMG_U8: return 1;
    case CUIMG_F32: return 4;
    case CUIMG_F32X4: return 16;
    case CUIMG_U32: return 4;
    case CUIMG_F32X2: return 8;
    default: return 0;
  }
}

// Launch bracket: records events around the launch when profiling is on and
// turns launch errors into error codes.
struct launch_scope {
  cuimg_hip_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  launch_scope(cuimg_hip_ctx* c, const char* n);
  int done();
};

template <typename F>
inline int launch(cuimg_hip_ctx* ctx, const char* name, F&& f) {
  launch_scope s(ctx, name);
  f();
  return s.done();
}

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

inline void add_child(cuimg_hip_ctx* ctx, void* p, void (*destroy)(void*)) { ctx->children.emplace_back(p, destroy); }
inline void drop_child(cuimg_hip_ctx* ctx, void* p) {
  for (size_t i = 0; i < ctx->children.size(); i++)
    if (ctx->children[i].first == p) {
      ctx->children.erase(ctx->children.begin() + (long)i);
      return;
    }
}

// Entry points shared between translation units.
int image_alloc(cuimg_hip_ctx* ctx, int nrows, int ncols, int border, int dtype, cuimg_img* out);
int image_free(cuimg_hip_ctx* ctx, cuimg_img* img);
int fill_border(cuimg_hip_ctx* ctx, const cuimg_img* img);
int convolve_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs, int n,
                       const float* coefs, int K, int th_override);
int gaussian_coefs(int sigma, int K, float* out);
int pyramid_build(cuimg_hip_ctx* ctx, const cuimg_img* frame, const cuimg_img* levels, int n);
int fast_saliency(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* mask, const cuimg_img* out,
                  int n, int th);

// ------------------------------------------------------------ device helpers
template <typename T>
__device__ __forceinline__ T* px(void* begin, int pitch, int r, int c) {
  return reinterpret_cast<T*>(static_cast<char*>(begin) + (ptrdiff_t)r * pitch) + c;
}
template <typename T>
__device__ __forceinline__ const T* px(const void* begin, int pitch, int r, int c) {
  return reinterpret_cast<const T*>(static_cast<const char*>(begin) + (ptrdiff_t)r * pitch) + c;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return max(lo, min(v, hi)); }  // lo <= hi

// Output size above which image kernels write with non-temporal stores: the 8
// XCDs' L2s (4 MiB each) could not hand such an image to the next kernel anyway.
constexpr size_t kStreamOutBytes = 32u << 20;

// Raw buffer view of an image: a load is one instruction addressed by a 32-bit
// VGPR byte offset, a wave-uniform SGPR byte offset and an immediate, with no
// 64-bit address arithmetic.  No range check (num_records = 2^32 - 1): callers
// keep offsets inside the image.  0x00020000 is the gfx9 family's raw 32-bit
// word format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
__device__ __forceinline__ float ld_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
}
__device__ __forceinline__ float2 ld_f32x2(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff) {
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0);
  return make_float2(__uint_as_float(w[0]), __uint_as_float(w[1]));
}

// Non-temporal stores (the line is not kept in L2).  Written in asm: as
// __builtin_nontemporal_store the compiler merges `nt ? nt_store : store` into
// one ordinary store.  Nothing in the kernels reads these locations back, so no
// wait is placed for them; the s_nop after a 16-byte store covers the gfx9
// hazard of overwriting a wide store's data registers.
typedef float v4f_t __attribute__((ext_vector_type(4)));
typedef float v2f_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(float* p, float v) {
  asm volatile("global_store_dword %0, %1, off nt" : : "v"(p), "v"(v) : "memory");
}
__device__ __forceinline__ void store_nt(uint8_t* p, uint8_t v) {
  asm volatile("global_store_byte %0, %1, off nt" : : "v"(p), "v"((uint32_t)v) : "memory");
}
__device__ __forceinline__ void store_nt(void* p, v2f_t v) {
  asm volatile("global_store_dwordx2 %0, %1, off nt" : : "v"(p), "v"(v) : "memory");
}
__device__ __forceinline__ void store_nt(void* p, v4f_t v) {
  asm volatile("global_store_dwordx4 %0, %1, off nt\n\ts_nop 1" : : "v"(p), "v"(v) : "memory");
}

// Calls f(integral_constant<int, 0>) ... f(integral_constant<int, N-1>) in order,
// so a loop body can use its index as a compile-time register index.
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on XCD
// b mod 8).  xcd_order(b, nb) renumbers a grid of nb workgroups so that each
// XCD takes a contiguous run of tile indices: neighbouring tiles, which share
// halo lines, then share an L2.
__device__ __forceinline__ int xcd_order(int b, int nb) {
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// A wave-uniform value moved to scalar registers (all lanes hold the same bits).
template <typename T>
__device__ __forceinline__ T uni(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "whole dwords");
  T r;
  const int* src = reinterpret_cast<const int*>(&v);
  int* dst = reinterpret_cast<int*>(&r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
  return r;
}

}  // namespace cuimg
