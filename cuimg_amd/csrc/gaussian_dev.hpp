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
  int stream_out;  // write the output with non-temporal stores
};

struct gauss_jobs {
  gauss_job j[kMaxJobs];
  int n;
};

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

// Vector store; nt: non-temporal (not kept in L2).
template <int VEC>
__device__ __forceinline__ void store_vec(float* p, const float (&v)[VEC], bool nt) {
  if constexpr (VEC == 4) {
    const v4f_t w = {v[0], v[1], v[2], v[3]};
    if (nt) store_nt(p, w);
    else *reinterpret_cast<v4f_t*>(p) = w;
  } else if constexpr (VEC == 2) {
    const v2f_t w = {v[0], v[1]};
    if (nt) store_nt(p, w);
    else *reinterpret_cast<v2f_t*>(p) = w;
  } else {
#pragma unroll
    for (int i = 0; i < VEC; i++) {
      if (nt) store_nt(p + i, v[i]);
      else p[i] = v[i];
    }
  }
}
template <int VEC>
__device__ __forceinline__ void store_vec(float4* p, const float4 (&v)[VEC], bool nt) {
#pragma unroll
  for (int i = 0; i < VEC; i++) {
    const v4f_t w = {v[i].x, v[i].y, v[i].z, v[i].w};
    if (nt) store_nt(p + i, w);
    else *reinterpret_cast<v4f_t*>(p + i) = w;
  }
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

// LDS of gauss_tile: two rows of W + 2K elements.
template <typename E, int VEC, int K>
constexpr int gauss_tile_lds_bytes() {
  return 2 * (64 * VEC + 2 * K) * (int)sizeof(E);
}

// One tile, staging rows in lds_mem (gauss_tile_lds_bytes, 16-byte aligned).
template <typename TI, typename E, int VEC, int K>
__device__ __forceinline__ void gauss_tile(const gauss_job& J, int tx, int ty, const coef_pack& cp,
                                           void* lds_mem) {
  constexpr int W = 64 * VEC;
  constexpr int P = 2 * K + 1;
  constexpr int LW = W + 2 * K;
  E (*lds)[LW] = reinterpret_cast<E (*)[LW]>(lds_mem);

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
          store_vec<VEC>(orow + col, o, J.stream_out);
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
__device__ __forceinline__ void gauss_tile(const gauss_job& J, int tx, int ty, const coef_pack& cp) {
  __shared__ __align__(16) char lds_mem[gauss_tile_lds_bytes<E, VEC, K>()];
  gauss_tile<TI, E, VEC, K>(J, tx, ty, cp, lds_mem);
}

// ------------------------------------------------------ direct-to-LDS path
// For full-width strips of F32 / F32X4 images the input rows stream straight
// from HBM into an LDS ring with global_load_lds (16 bytes per lane, no
// registers held while in flight), D rows ahead of the row being filtered, so
// every wave keeps D rows of memory traffic in flight instead of one.  Each
// ring row is [left halo K][strip W][right halo K] with the strip 16-byte
// aligned; a row costs three DMA instructions (strip, left halo, right halo).
//
// Waits are placed by hand.  Loads complete in issue order, so while any DMA
// of row r is pending, the 3 (D - 1) DMAs of the rows issued after it are too:
// at least 3 (D - 1) + 1 operations are outstanding.  Waiting until at most
// 3 (D - 1) are outstanding therefore guarantees row r has landed, however the
// stores issued in between complete.  Each step stores the previous step's
// output right after its wait.
//
// Odd tile rows stream bottom-up: two vertically adjacent tiles then read the
// 2K rows they share at about the same time (both first or both last), so the
// second read is an L2 hit when the tiles run on one XCD (the launch orders
// tiles per XCD, gaussian.hip).  Outputs too large to stay in cache are
// written with non-temporal stores (gauss_job::stream_out).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(3))) void* lds_addr(T* p) {
  return (__attribute__((address_space(3))) void*)p;
}

template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) void* glb_addr(const T* p) {
  return (__attribute__((address_space(1))) void*)(p);
}

// LDS reads of DMA-filled rows.  The compiler cannot tell which DMA wrote
// which ring slot and would wait for all of them before any LDS read, so these
// reads are written as ds_read_b128 in asm, each block waiting for its own
// results (lgkmcnt(0)) before it ends; the DMA waits are placed by hand.
__device__ __forceinline__ uint32_t lds_u32(const void* p) {
  return (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const void*)p;
}

template <int N>
__device__ __forceinline__ void lds_read16(const uint32_t* a, float4* r);

template <>
__device__ __forceinline__ void lds_read16<1>(const uint32_t* a, float4* r) {
  asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=&v"(r[0]) : "v"(a[0]) : "memory");
}
template <>
__device__ __forceinline__ void lds_read16<2>(const uint32_t* a, float4* r) {
  asm volatile("ds_read_b128 %0, %2\n\tds_read_b128 %1, %3\n\ts_waitcnt lgkmcnt(0)"
               : "=&v"(r[0]), "=&v"(r[1])
               : "v"(a[0]), "v"(a[1])
               : "memory");
}
template <>
__device__ __forceinline__ void lds_read16<3>(const uint32_t* a, float4* r) {
  asm volatile("ds_read_b128 %0, %3\n\tds_read_b128 %1, %4\n\tds_read_b128 %2, %5\n\ts_waitcnt lgkmcnt(0)"
               : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2])
               : "v"(a[0]), "v"(a[1]), "v"(a[2])
               : "memory");
}
template <>
__device__ __forceinline__ void lds_read16<4>(const uint32_t* a, float4* r) {
  asm volatile(
      "ds_read_b128 %0, %4\n\tds_read_b128 %1, %5\n\tds_read_b128 %2, %6\n\tds_read_b128 %3, %7\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3])
      : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3])
      : "memory");
}

// 16-byte store out of the compiler's view, so that no wait for it is placed
// (nothing in the kernel reads what it writes); the s_nop covers the gfx9
// hazard of overwriting a wide store's data registers.  nt: non-temporal.
__device__ __forceinline__ void global_store16(void* p, float4 v, bool nt) {
  const v4f_t w = {v.x, v.y, v.z, v.w};
  if (nt) store_nt(p, w);
  else asm volatile("global_store_dwordx4 %0, %1, off\n\ts_nop 1" : : "v"(p), "v"(w) : "memory");
}

// r[k] = the 16 bytes at LDS byte address base + 16 k, k < N, in blocks of 4.
template <int N>
__device__ __forceinline__ void lds_read16_run(uint32_t base, float4 (&r)[N]) {
  static_for<(N + 3) / 4>([&](auto gc) {
    constexpr int g = decltype(gc)::value;
    constexpr int cnt = N - 4 * g < 4 ? N - 4 * g : 4;
    uint32_t a[4];
#pragma unroll
    for (int k = 0; k < cnt; k++) a[k] = base + 16u * (uint32_t)(4 * g + k);
    lds_read16<cnt>(a, &r[4 * g]);
  });
}

template <int N>
__device__ __forceinline__ void wait_vm() {
  // s_waitcnt vmcnt(N), expcnt and lgkmcnt left at their maxima (gfx9 encoding)
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Ring geometry: R = D + 1 rows of RW elements, [pad HL][strip W][halo K].
template <typename E, int VEC, int K, int D>
struct dma_ring {
  static constexpr int W = 64 * VEC;
  static constexpr int R = D + 1;
  static constexpr int EPB = 16 / (int)sizeof(E);       // elements per 16 bytes
  static constexpr int HL = (K + EPB - 1) / EPB * EPB;  // left pad: strip 16-byte aligned
  static constexpr int RW = (HL + W + K + EPB - 1) / EPB * EPB;
  static constexpr int bytes = R * RW * (int)sizeof(E);
};

// One tile (UP: bottom-up), its ring in ring_mem (dma_ring::bytes, 16-byte
// aligned; both directions of a kernel share it).
template <typename E, int VEC, int K, int D, bool UP>
__device__ __forceinline__ void gauss_tile_dma(const gauss_job& J, int tx, int ty, const coef_pack& cp,
                                               void* ring_mem) {
  using G = dma_ring<E, VEC, K, D>;
  constexpr int W = G::W;
  constexpr int P = 2 * K + 1;
  constexpr int R = G::R;
  constexpr int EPB = G::EPB;
  constexpr int HL = G::HL;
  constexpr int RW = G::RW;
  static_assert(sizeof(E) * VEC == 16, "16 bytes per lane");
  static_assert(D >= 2 && 3 * (D - 1) < 64, "vmcnt range");
  constexpr unsigned ESZ = (unsigned)sizeof(E);  // DMA size operands must be 32-bit constants
  E (*ring)[RW] = reinterpret_cast<E (*)[RW]>(ring_mem);

  const int lane = threadIdx.x;
  const int c0 = tx * W;
  const int r0 = ty * J.th;
  const int r_end = min(r0 + J.th, J.nr);
  const int first = r0 - K, last = r_end - 1 + K;
  const int nsteps = last - first + 1;  // step q filters input row row_of(q)
  const int col = c0 + VEC * lane;
  const int hcl = max(c0 - K + lane, 0), hcr = min(c0 + W + lane, J.nc - 1);
  const bool nt = J.stream_out;

  float c[P];
#pragma unroll
  for (int m = 0; m < P; m++) c[m] = cp.c[m];

  auto row_of = [&](int q) { return UP ? last - q : first + q; };
  auto issue = [&](int q) {
    E* row_lds = ring[__builtin_amdgcn_readfirstlane(q % R)];
    const E* row = px<E>(J.in, J.in_pitch, clampi(row_of(q), 0, J.nr - 1), 0);
    __builtin_amdgcn_global_load_lds(glb_addr(row + col), lds_addr(row_lds + HL), 16u, 0, 0);
    if (lane < K) {
      __builtin_amdgcn_global_load_lds(glb_addr(row + hcl), lds_addr(row_lds + HL - K), ESZ, 0, 0);
      __builtin_amdgcn_global_load_lds(glb_addr(row + hcr), lds_addr(row_lds + HL + W), ESZ, 0, 0);
    }
  };
#pragma unroll
  for (int d = 0; d < D; d++)
    if (d < nsteps) issue(d);

  E hist[P][VEC];
  E pend[VEC];        // output row of the previous step, stored after this step's wait
  int pend_row = -1;
  for (int qb = 0; qb < nsteps; qb += P) {
    static_for<P>([&](auto tc) {
      constexpr int t = decltype(tc)::value;
      const int q = qb + t;
      if (q >= nsteps) return;
      if (q + D - 1 >= nsteps) wait_vm<0>();
      else wait_vm<3 * (D - 1)>();
      if (pend_row >= 0)
        global_store16(px<E>(J.out, J.out_pitch, pend_row, 0) + col, *reinterpret_cast<float4*>(pend), nt);
      pend_row = -1;
      const E* row_lds = ring[q % R];
      // 16-byte chunks covering elements [HL - K + VEC*lane, HL + K + VEC*(lane+1))
      constexpr int CL = (HL - K) / EPB;
      constexpr int NCH = (HL + K + VEC + EPB - 1) / EPB - CL;
      float4 ch[NCH];
      lds_read16_run<NCH>(lds_u32(row_lds) + 16u * (uint32_t)(CL + lane * VEC / EPB), ch);
      E x[VEC + 2 * K];
      const E* che = reinterpret_cast<const E*>(ch);
#pragma unroll
      for (int m = 0; m < VEC + 2 * K; m++) x[m] = che[HL - K - CL * EPB + m];
      // the slot of step q + D held step q - 1, read by the previous step
      if (q + D < nsteps) issue(q + D);
#pragma unroll
      for (int i = 0; i < VEC; i++) {
        E s = mul_e(c[2 * K], x[i + 2 * K]);
#pragma unroll
        for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], x[i + m], s);
        hist[t][i] = s;
      }
      if (q >= 2 * K) {
        // output row o = row_of(q) -+ K; tap m (weight c[m]) is row o - K + m,
        // accumulated from m = 2K down to 0.  Going down, row o - K + m was
        // filtered at step q - 2K + m (slot (t + m + 1) mod P); going up, at
        // step q - m (slot (t + P - m) mod P).
#pragma unroll
        for (int i = 0; i < VEC; i++) {
          E s;
          if constexpr (!UP) {
            s = mul_e(c[2 * K], hist[t][i]);
#pragma unroll
            for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], hist[(t + m + 1) % P][i], s);
          } else {
            s = mul_e(c[2 * K], hist[(t + 1) % P][i]);
#pragma unroll
            for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], hist[(t + P - m) % P][i], s);
          }
          pend[i] = s;
        }
        pend_row = UP ? row_of(q) + K : row_of(q) - K;
      }
    });
  }
  if (pend_row >= 0)
    global_store16(px<E>(J.out, J.out_pitch, pend_row, 0) + col, *reinterpret_cast<float4*>(pend), nt);
}

}  // namespace cuimg
