// Separable Gaussian for gfx950: one fused pass per image, no temporary image.
//
// Semantics follow the reference's meta-template GPU kernels
// (cuimg/gpu/meta_convolve.h:25-128): a row pass then a column pass over taps
// x = -K..K with clamp-to-edge reads, FP32 arithmetic, each sum accumulated from
// tap +K down to tap -K (the association of the recursive template).
//
// Design (HBM-bound stencil): one wave owns a strip of 64*VEC columns and
// TH output rows.  It streams input rows top to bottom: each row is loaded once
// with coalesced vector loads (plus K halo columns per side), staged in LDS,
// filtered horizontally from LDS into a register ring that holds the last 2K+1
// filtered rows, and the vertical tap sum of the ring gives one output row,
// stored with vector stores.  The row loop is unrolled by the ring length so
// ring slots are compile-time registers.  The next row's global loads are
// issued before the current row is filtered.  Input traffic is 1x per strip
// plus the 2K-row warm-up per tile; the intermediate never leaves registers.
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <utility>

namespace cuimg {

constexpr int kMaxK = 16;
constexpr int kMaxJobs = 8;

struct coef_pack {
  float c[2 * kMaxK + 1];
};

struct gauss_job {
  const void* in;
  void* out;
  int in_pitch, out_pitch;
  int nr, nc;
  int th;          // output rows per tile
  int tiles_x;
  int tile_begin;  // first flattened tile index of this job
  int vec_ok;      // in/out aligned for vector access
};

struct gauss_jobs {
  gauss_job j[kMaxJobs];
  int n;
};

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

// ---------------------------------------------------------------- elements
__device__ __forceinline__ float fma_e(float a, float x, float s) { return fmaf(a, x, s); }
__device__ __forceinline__ float4 fma_e(float a, float4 x, float4 s) {
  return make_float4(fmaf(a, x.x, s.x), fmaf(a, x.y, s.y), fmaf(a, x.z, s.z), fmaf(a, x.w, s.w));
}
__device__ __forceinline__ float mul_e(float a, float x) { return a * x; }
__device__ __forceinline__ float4 mul_e(float a, float4 x) {
  return make_float4(a * x.x, a * x.y, a * x.z, a * x.w);
}

template <typename TI, typename E>
struct io;

template <>
struct io<uint8_t, float> {
  template <int VEC>
  __device__ static void load(const uint8_t* p, float (&v)[VEC]) {
    if constexpr (VEC == 4) {
      uchar4 q = *reinterpret_cast<const uchar4*>(p);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
#pragma unroll
      for (int i = 0; i < VEC; i++) v[i] = p[i];
    }
  }
  __device__ static float load1(const uint8_t* p) { return (float)*p; }
};

template <>
struct io<float, float> {
  template <int VEC>
  __device__ static void load(const float* p, float (&v)[VEC]) {
    if constexpr (VEC == 4) {
      float4 q = *reinterpret_cast<const float4*>(p);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else if constexpr (VEC == 2) {
      float2 q = *reinterpret_cast<const float2*>(p);
      v[0] = q.x; v[1] = q.y;
    } else {
#pragma unroll
      for (int i = 0; i < VEC; i++) v[i] = p[i];
    }
  }
  __device__ static float load1(const float* p) { return *p; }
};

template <>
struct io<float4, float4> {
  template <int VEC>
  __device__ static void load(const float4* p, float4 (&v)[VEC]) {
#pragma unroll
    for (int i = 0; i < VEC; i++) v[i] = p[i];
  }
  __device__ static float4 load1(const float4* p) { return *p; }
};

template <int VEC>
__device__ __forceinline__ void store_vec(float* p, const float (&v)[VEC]) {
  if constexpr (VEC == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (VEC == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int i = 0; i < VEC; i++) p[i] = v[i];
  }
}
template <int VEC>
__device__ __forceinline__ void store_vec(float4* p, const float4 (&v)[VEC]) {
#pragma unroll
  for (int i = 0; i < VEC; i++) p[i] = v[i];
}

// Loads the VEC elements of this lane plus the halo elements held by lanes < K.
template <typename TI, typename E, int VEC, int K>
__device__ __forceinline__ void load_row(const gauss_job& J, int ri, int c0, int lane, E (&v)[VEC],
                                         E& hl, E& hr) {
  constexpr int W = 64 * VEC;
  const int rc = clampi(ri, 0, J.nr - 1);
  const TI* row = px<TI>(J.in, J.in_pitch, rc, 0);
  const int col = c0 + VEC * lane;
  if (J.vec_ok && col + VEC <= J.nc) {
    io<TI, E>::template load<VEC>(row + col, v);
  } else {
#pragma unroll
    for (int i = 0; i < VEC; i++) v[i] = io<TI, E>::load1(row + min(col + i, J.nc - 1));
  }
  if (lane < K) {
    hl = io<TI, E>::load1(row + max(c0 - K + lane, 0));
    hr = io<TI, E>::load1(row + min(c0 + W + lane, J.nc - 1));
  }
}

template <typename TI, typename E, int VEC, int K>
__device__ __forceinline__ void gauss_tile(const gauss_job& J, int tx, int ty, const coef_pack& cp) {
  constexpr int W = 64 * VEC;
  constexpr int P = 2 * K + 1;
  constexpr int LW = W + 2 * K;
  __shared__ E lds[2][LW];

  const int lane = threadIdx.x;
  const int c0 = tx * W;
  const int r0 = ty * J.th;
  const int r_end = min(r0 + J.th, J.nr);
  const int first = r0 - K, last = r_end - 1 + K;
  const int col = c0 + VEC * lane;

  float c[P];
#pragma unroll
  for (int m = 0; m < P; m++) c[m] = cp.c[m];

  E hist[P][VEC];
  E cur[VEC], hl = E(), hr = E();
  load_row<TI, E, VEC, K>(J, first, c0, lane, cur, hl, hr);
  int buf = 0;

  for (int base = first; base <= last; base += P) {
    // ring slot t holds row base + t; the P steps are expanded at compile time
    static_for<P>([&](auto tc) {
      constexpr int t = decltype(tc)::value;
      const int ri = base + t;
      if (ri > last) return;
#pragma unroll
      for (int i = 0; i < VEC; i++) lds[buf][K + VEC * lane + i] = cur[i];
      if (lane < K) {
        lds[buf][lane] = hl;
        lds[buf][K + W + lane] = hr;
      }
      if (ri + 1 <= last) load_row<TI, E, VEC, K>(J, ri + 1, c0, lane, cur, hl, hr);
      __syncthreads();

      E x[VEC + 2 * K];
#pragma unroll
      for (int m = 0; m < VEC + 2 * K; m++) x[m] = lds[buf][VEC * lane + m];
#pragma unroll
      for (int i = 0; i < VEC; i++) {
        E s = mul_e(c[2 * K], x[i + 2 * K]);
#pragma unroll
        for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], x[i + m], s);
        hist[t][i] = s;
      }
      buf ^= 1;

      if (ri - K >= r0) {
        // rows ri-2K .. ri live in slots (t - 2K + m) mod P = (t + m + 1) mod P
        E o[VEC];
#pragma unroll
        for (int i = 0; i < VEC; i++) {
          E s = mul_e(c[2 * K], hist[t][i]);
#pragma unroll
          for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], hist[(t + m + 1) % P][i], s);
          o[i] = s;
        }
        E* orow = px<E>(J.out, J.out_pitch, ri - K, 0);
        if (J.vec_ok && col + VEC <= J.nc) {
          store_vec<VEC>(orow + col, o);
        } else {
#pragma unroll
          for (int i = 0; i < VEC; i++)
            if (col + i < J.nc) orow[col + i] = o[i];
        }
      }
    });
  }
}

}  // namespace cuimg
