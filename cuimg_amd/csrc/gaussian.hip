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

template <typename TI, typename E, int VEC, int K>
__global__ __launch_bounds__(64) void gauss_sep_kernel(gauss_jobs jobs, coef_pack cp) {
  int b = blockIdx.x;
  int k = 0;
  while (k + 1 < jobs.n && b >= jobs.j[k + 1].tile_begin) k++;
  const gauss_job& J = jobs.j[k];
  b -= J.tile_begin;
  gauss_tile<TI, E, VEC, K>(J, b % J.tiles_x, b / J.tiles_x, cp);
}

// ------------------------------------------------------------------ host
template <typename TI, typename E, int VEC, int K>
static void launch_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp) {
  gauss_sep_kernel<TI, E, VEC, K><<<blocks, 64, 0, ctx->stream>>>(jobs, cp);
}

template <typename TI, typename E, int VEC>
static int dispatch_k(cuimg_hip_ctx* ctx, const gauss_jobs& jobs, int blocks, const coef_pack& cp, int K) {
  switch (K) {
#define CUIMG_K(k) case k: launch_k<TI, E, VEC, k>(ctx, jobs, blocks, cp); return 0;
    CUIMG_K(1) CUIMG_K(2) CUIMG_K(3) CUIMG_K(4) CUIMG_K(5) CUIMG_K(6) CUIMG_K(7) CUIMG_K(8)
    CUIMG_K(9) CUIMG_K(10) CUIMG_K(11) CUIMG_K(12) CUIMG_K(13) CUIMG_K(14) CUIMG_K(15) CUIMG_K(16)
#undef CUIMG_K
  }
  return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported (1..%d)", K, kMaxK);
}

static int pick_th(int nr, int tiles_x, int total_tiles_other) {
  // aim for >= ~2048 waves in flight over the whole launch, 16..128 rows per tile
  long want = 2048 - total_tiles_other;
  if (want < 256) want = 256;
  long th = ((long)nr * tiles_x + want - 1) / want;
  th = std::max<long>(16, std::min<long>(128, th));
  return (int)th;
}

int convolve_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs, int n,
                       const float* coefs, int K, int th_override) {
  if (n <= 0 || n > kMaxJobs || !ins || !outs || !coefs)
    return set_error(ctx, CUIMG_E_INVALID, "gaussian: bad job list (n=%d)", n);
  if (K < 1 || K > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported (1..%d)", K, kMaxK);
  const int in_dt = ins[0].dtype;
  const int vec = in_dt == CUIMG_F32X4 ? 1 : 4;
  const int W = 64 * vec;
  const int es_in = dtype_size(in_dt), es_out = in_dt == CUIMG_F32X4 ? 16 : 4;
  gauss_jobs jobs{};
  jobs.n = n;
  int total = 0;
  for (int i = 0; i < n; i++) {
    const cuimg_img& a = ins[i];
    const cuimg_img& b = outs[i];
    if (a.dtype != in_dt || !(in_dt == CUIMG_U8 || in_dt == CUIMG_F32 || in_dt == CUIMG_F32X4))
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: input dtype %d not supported", a.dtype);
    const int want_out = in_dt == CUIMG_F32X4 ? CUIMG_F32X4 : CUIMG_F32;
    if (b.dtype != want_out)
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: output dtype %d, expected %d", b.dtype, want_out);
    if (a.nrows != b.nrows || a.ncols != b.ncols || a.nrows <= 0 || a.ncols <= 0 || !a.begin || !b.begin)
      return set_error(ctx, CUIMG_E_INVALID, "gaussian: domains differ or empty (%dx%d vs %dx%d)",
                       a.nrows, a.ncols, b.nrows, b.ncols);
    gauss_job& J = jobs.j[i];
    J.in = a.begin;
    J.out = b.begin;
    J.in_pitch = a.pitch;
    J.out_pitch = b.pitch;
    J.nr = a.nrows;
    J.nc = a.ncols;
    J.tiles_x = ceil_div(a.ncols, W);
    J.vec_ok = ((uintptr_t)a.begin % (vec * es_in) == 0) && (a.pitch % (vec * es_in) == 0) &&
               ((uintptr_t)b.begin % (vec * es_out) == 0) && (b.pitch % (vec * es_out) == 0);
  }
  // tile height: sized on the largest image
  int big = 0;
  for (int i = 1; i < n; i++)
    if ((long)jobs.j[i].nr * jobs.j[i].nc > (long)jobs.j[big].nr * jobs.j[big].nc) big = i;
  const int th = th_override > 0 ? th_override : pick_th(jobs.j[big].nr, jobs.j[big].tiles_x, 0);
  for (int i = 0; i < n; i++) {
    gauss_job& J = jobs.j[i];
    J.th = th;
    J.tile_begin = total;
    total += J.tiles_x * ceil_div(J.nr, th);
  }
  coef_pack cp{};
  for (int m = 0; m < 2 * K + 1; m++) cp.c[m] = coefs[m];
  return launch(ctx, "gaussian_sep", [&] {
    int r = 0;
    if (in_dt == CUIMG_U8) r = dispatch_k<uint8_t, float, 4>(ctx, jobs, total, cp, K);
    else if (in_dt == CUIMG_F32) r = dispatch_k<float, float, 4>(ctx, jobs, total, cp, K);
    else r = dispatch_k<float4, float4, 1>(ctx, jobs, total, cp, K);
    (void)r;
  });
}

// gaussian_derivative<0> coefficients as emitted by the reference's generator:
// the double-precision formula rounded to float, printed with 6 significant
// digits and read back as a float literal
// (tools/make_meta_gaussian/make_meta_gaussian.cpp:11-16, 85-89).
int gaussian_coefs(int sigma, int K, float* out) {
  if (sigma < 1 || K < 0 || !out) return CUIMG_E_INVALID;
  const float pi = 3.14159265f;  // CUIMG_PI
  const float s = (float)sigma;
  for (int x = -K; x <= K; x++) {
    double a = 1 / (std::sqrt((double)(2 * pi)) * (double)s);
    float e = (float)(-(x * x)) / (2 * s * s);
    float v = (float)(a * std::exp((double)e));
    char buf[64];
    std::snprintf(buf, sizeof buf, "%g", (double)v);
    out[x + K] = std::strtof(buf, nullptr);
  }
  return 0;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_gaussian_coefs(int sigma, int half_width, float* out) {
  return gaussian_coefs(sigma, half_width, out);
}

int cuimg_hip_convolve_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           const float* coefs, int half_width) {
  return convolve_sep_multi(ctx, in, out, 1, coefs, half_width, 0);
}

int cuimg_hip_gaussian_sep(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                           int sigma, int half_width) {
  float c[2 * kMaxK + 1];
  if (half_width < 1 || half_width > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported", half_width);
  CUIMG_TRY(gaussian_coefs(sigma, half_width, c));
  return convolve_sep_multi(ctx, in, out, 1, c, half_width, 0);
}

int cuimg_hip_gaussian_sep_multi(cuimg_hip_ctx* ctx, const cuimg_img* ins, const cuimg_img* outs,
                                 int n, int sigma, int half_width) {
  float c[2 * kMaxK + 1];
  if (half_width < 1 || half_width > kMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "gaussian: half width %d not supported", half_width);
  CUIMG_TRY(gaussian_coefs(sigma, half_width, c));
  return convolve_sep_multi(ctx, ins, outs, n, c, half_width, 0);
}

// Tile-height override for the autotuner and the roofline sweep.
int cuimg_hip_convolve_sep_th(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out,
                              const float* coefs, int half_width, int th) {
  return convolve_sep_multi(ctx, in, out, 1, coefs, half_width, th);
}

}  // extern "C"
