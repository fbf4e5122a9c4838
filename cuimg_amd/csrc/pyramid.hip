// Box pyramid: standalone launch (the device code and its reference notes are
// in pyramid_dev.hpp), and the blur-then-decimate level.
#include "gaussian_dev.hpp"
#include "pyramid_dev.hpp"

#include <algorithm>
#include <type_traits>

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

// ------------------------------------------------- Gaussian then decimate
// The blur-and-subsample pyramid of the reference's disabled build_mipmap
// (cuimg/gpu/mipmap.h:94-109): local_jet_static<..., 0, 0, 1, 6> (sigma 1,
// K = 6, clamp-to-edge, row pass then column pass, each sum from tap +K down to
// tap -K) and the sample at 2p, out domain = in / 2 (floor).
//
// Block kernel (launched for U8 input; F32 images take the streamed kernel
// below, measured 1.7x faster), fused: a tile is 16 x 64 output pixels.  Its
// (2*16 - 1 + 12)-row input block is staged in LDS (columns 2*c0 - 8 ..
// 2*c0 + 136), rows are filtered at the even columns only, then
// the even rows are filtered: no full-resolution intermediate reaches memory.
// A workgroup walks down a run of kDecRun tiles of one tile column, loading the
// next tile's block into registers while it filters the current one, so HBM
// stays busy through the LDS phases; consecutive runs of a row of runs go to
// one XCD, so horizontally adjacent tiles read their shared halo columns from
// one L2, and a run's vertical halo rows were read by the run itself.
constexpr int kDecK = 6, kDecTH = 16, kDecTW = 64, kDecRun = 4;
constexpr int kDecInR = 2 * kDecTH - 1 + 2 * kDecK;  // input rows of a tile
constexpr int kDecInC = 2 * kDecTW + 16;              // input columns, 16-byte aligned span
constexpr int kDecQ = kDecInC / 4;                    // 16-byte chunks per input row
constexpr int kDecInStride = kDecInC + 4;
constexpr int kDecChunks = (kDecInR * kDecQ + 255) / 256;  // per thread

struct dec_coefs {
  float c[2 * kDecK + 1];
};

struct dec_args {
  const void* in;
  int in_pitch, nr, nc;
  float* out;
  int out_pitch, onr, onc;
  int tiles_x, tiles_y, runs_y;
  int vec_ok;      // input 16-byte aligned (rows and start)
  int out_vec_ok;  // output 8-byte aligned
  dec_coefs cf;
};

// Chunk k of this thread in the staged block of the tile whose input origin is
// (ir0, ic0); rows and columns clamp to the domain.  Addresses are 32-bit byte
// offsets from the image start (the image spans less than 4 GiB), so a load
// holds one offset register next to the scalar base.
//
// dec_chunk_vec (aligned F32 image, ic0 a multiple of 4): one 16-byte load of
// the chunk, or for a chunk reaching past either side, of the in-domain chunk
// holding its clamped columns, from which each element is picked.
__device__ __forceinline__ float4 dec_chunk_vec(const dec_args& a, int ir0, int ic0, int k, bool edge) {
  const int i = k * 256 + threadIdx.x;
  const int rr = i / kDecQ, q = i % kDecQ;
  const uint32_t row = (uint32_t)clampi(ir0 + rr, 0, a.nr - 1) * (uint32_t)a.in_pitch;
  const int c = ic0 + 4 * q;
  const char* base = static_cast<const char*>(a.in);
  if (!edge) return *reinterpret_cast<const float4*>(base + row + 4u * (uint32_t)c);
  const int m = a.nc - 1, cl = clampi(c, 0, m & ~3);
  const float4 v = *reinterpret_cast<const float4*>(base + row + 4u * (uint32_t)cl);
  auto pick = [&](int j) {
    const int e = clampi(c + j, 0, m) - cl;
    return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w));
  };
  return make_float4(pick(0), pick(1), pick(2), pick(3));
}

// dec_chunk_clamped (any image): four column-clamped loads.
template <typename TI>
__device__ __forceinline__ float4 dec_chunk_clamped(const dec_args& a, int ir0, int ic0, int k) {
  const int i = k * 256 + threadIdx.x;
  const int rr = i / kDecQ, q = i % kDecQ;
  const int c = ic0 + 4 * q, m = a.nc - 1;
  const uint32_t row = (uint32_t)clampi(ir0 + rr, 0, a.nr - 1) * (uint32_t)a.in_pitch;
  const char* base = static_cast<const char*>(a.in);
  auto at = [&](int cc) {
    return (float)*reinterpret_cast<const TI*>(base + row + (uint32_t)clampi(cc, 0, m) * (uint32_t)sizeof(TI));
  };
  return make_float4(at(c), at(c + 1), at(c + 2), at(c + 3));
}

__device__ __forceinline__ float4* dec_dst(float* tin, int k) {
  const int i = k * 256 + threadIdx.x;
  return reinterpret_cast<float4*>(&tin[(i / kDecQ) * kDecInStride + 4 * (i % kDecQ)]);
}

// The next tile's block is held in seven float4 registers p0..p6 of the
// kernel, written out chunk by chunk (kept in an array or a struct, the block
// went to scratch memory).
#define CUIMG_DEC_CHUNKS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6)
static_assert(kDecChunks == 7, "seven chunks per thread");

template <typename TI, bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void gauss_decimate_kernel(dec_args a) {
  __shared__ __align__(16) float tin[kDecInR * kDecInStride];
  __shared__ float trow[kDecInR * kDecTW];
  const int run = xcd_order(blockIdx.x, gridDim.x);
  const int tx = run % a.tiles_x, ty0 = (run / a.tiles_x) * kDecRun;
  const int ty1 = min(ty0 + kDecRun, a.tiles_y);
  const int c0 = tx * kDecTW;
  float4 p0, p1, p2, p3, p4, p5, p6;
  const bool last_ok = 6 * 256 + (int)threadIdx.x < kDecInR * kDecQ;
  // VEC (aligned F32 image): 16-byte chunks, and the next tile's block is
  // loaded during the filters; otherwise clamped loads when a block is due
  const bool edge = 2 * c0 - 8 < 0 || 2 * c0 - 8 + kDecInC > a.nc;  // the same for the whole run
#define CUIMG_DEC_LOAD_VEC(k) \
  if (k < 6 || last_ok) p##k = dec_chunk_vec(a, 2 * (r0n) - kDecK, 2 * c0 - 8, k, edge);
#define CUIMG_DEC_LOAD_CLAMPED(k) \
  if (k < 6 || last_ok) p##k = dec_chunk_clamped<TI>(a, 2 * (r0n) - kDecK, 2 * c0 - 8, k);
#define CUIMG_DEC_PUT(k) \
  if (k < 6 || last_ok) *dec_dst(tin, k) = p##k;
  if constexpr (VEC) {
    const int r0n = ty0 * kDecTH;
    CUIMG_DEC_CHUNKS(CUIMG_DEC_LOAD_VEC)
  }
  for (int ty = ty0; ty < ty1; ty++) {
    const int r0 = ty * kDecTH;
    if constexpr (!VEC) {
      const int r0n = r0;
      CUIMG_DEC_CHUNKS(CUIMG_DEC_LOAD_CLAMPED)
    }
    CUIMG_DEC_CHUNKS(CUIMG_DEC_PUT)
    __syncthreads();
    if constexpr (VEC) {
      if (ty + 1 < ty1) {
        const int r0n = r0 + kDecTH;
        CUIMG_DEC_CHUNKS(CUIMG_DEC_LOAD_VEC)
      }
    }
    // row pass: a thread filters two adjacent even columns 2*oc, 2*oc + 2
    // (staged columns 2*oc + 8 and 2*oc + 10) from five 16-byte LDS reads
    for (int i = threadIdx.x; i < kDecInR * (kDecTW / 2); i += 256) {
      const int rr = i / (kDecTW / 2), j = i % (kDecTW / 2);
      const float4* t4 = reinterpret_cast<const float4*>(&tin[rr * kDecInStride + 4 * j]);
      float w[20];
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const float4 v = t4[q];
        w[4 * q] = v.x;
        w[4 * q + 1] = v.y;
        w[4 * q + 2] = v.z;
        w[4 * q + 3] = v.w;
      }
      // output column oc = 2j reads staged 4j + 2 .. 4j + 14, oc = 2j + 1 reads 4j + 4 .. 4j + 16
      float acc0 = a.cf.c[2 * kDecK] * w[2 + 2 * kDecK];
      float acc1 = a.cf.c[2 * kDecK] * w[4 + 2 * kDecK];
#pragma unroll
      for (int k = 2 * kDecK - 1; k >= 0; k--) {
        acc0 = fmaf(a.cf.c[k], w[2 + k], acc0);
        acc1 = fmaf(a.cf.c[k], w[4 + k], acc1);
      }
      trow[rr * kDecTW + 2 * j] = acc0;
      trow[rr * kDecTW + 2 * j + 1] = acc1;
    }
    __syncthreads();
    // column pass: a thread filters output rows 2m and 2m + 1 of one column
    for (int i = threadIdx.x; i < (kDecTH / 2) * kDecTW; i += 256) {
      const int m = i / kDecTW, oc = i % kDecTW;
      const float* t = &trow[4 * m * kDecTW + oc];
      float w[2 * kDecK + 3];
#pragma unroll
      for (int k = 0; k < 2 * kDecK + 3; k++) w[k] = t[k * kDecTW];
      float acc0 = a.cf.c[2 * kDecK] * w[2 * kDecK];
      float acc1 = a.cf.c[2 * kDecK] * w[2 * kDecK + 2];
#pragma unroll
      for (int k = 2 * kDecK - 1; k >= 0; k--) {
        acc0 = fmaf(a.cf.c[k], w[k], acc0);
        acc1 = fmaf(a.cf.c[k], w[k + 2], acc1);
      }
      const int r = r0 + 2 * m, c = c0 + oc;
      if (c < a.onc) {
        if (r < a.onr) *px<float>(a.out, a.out_pitch, r, c) = acc0;
        if (r + 1 < a.onr) *px<float>(a.out, a.out_pitch, r + 1, c) = acc1;
      }
    }
  }
}
#undef CUIMG_DEC_LOAD_VEC
#undef CUIMG_DEC_LOAD_CLAMPED
#undef CUIMG_DEC_PUT

// F32 input: one wave per strip of 256 input columns (128 output columns, two
// per lane) and kDecSTH output rows.  The wave streams the strip's input rows
// top to bottom into an LDS ring with global_load_lds, kDecD rows ahead (no
// registers held while they are in flight), filters every input row at its
// even columns into a register ring of the last 2K+1 filtered rows, and on
// every other row (input row 2o + K) sums the ring into output row o.  The
// arithmetic is the block kernel's: the same taps in the same order.  Strips
// inside aligned images load each row as one 16-byte access per lane; the last
// (partial) strip and unaligned images load 4 bytes per lane with clamped
// columns.  Waits are counted as in gauss_tile_dma: loads complete in order,
// so at most OPS (kDecD - 1) outstanding operations means the row has landed.
// Odd tile rows stream bottom-up (UP), so vertically adjacent tiles read the
// input rows they share at about the same time, from one XCD's L2.
constexpr int kDecSW = 256;                     // input columns per strip
constexpr int kDecSTH = 16;                     // output rows per tile
constexpr int kDecD = 4;                        // input rows in flight
constexpr int kDecHL = 8;                       // left pad: strip 16-byte aligned in LDS
constexpr int kDecRW = kDecHL + kDecSW + 8;     // ring row: pad, strip, right halo
constexpr int kDecR = kDecD + 1;                // ring rows

template <bool WIDE, bool UP>
__device__ __forceinline__ void dec_tile_dma(const dec_args& a, int tx, int ty, float* ring_mem) {
  constexpr int K = kDecK, P = 2 * K + 1;
  constexpr int OPS = WIDE ? 3 : 6;  // DMA instructions per row: strip (1 or 4) and two halos
  static_assert(OPS * (kDecD - 1) < 64, "vmcnt range");
  static_assert(kDecHL >= K && kDecRW >= kDecHL + kDecSW + K, "halos fit the ring row");
  float (*ring)[kDecRW] = reinterpret_cast<float (*)[kDecRW]>(ring_mem);
  const int lane = threadIdx.x;
  const int ic0 = tx * kDecSW;            // first input column of the strip
  const int oc = ic0 / 2 + 2 * lane;      // this lane's output columns oc, oc + 1
  const int o0 = ty * kDecSTH, o1 = min(o0 + kDecSTH, a.onr);
  const int first = 2 * o0 - K, last = 2 * (o1 - 1) + K;
  const int nsteps = last - first + 1;    // step q stages input row row_of(q)
  const int hcl = max(ic0 - K + lane, 0), hcr = min(ic0 + kDecSW + lane, a.nc - 1);

  float c[P];
#pragma unroll
  for (int m = 0; m < P; m++) c[m] = a.cf.c[m];

  auto row_of = [&](int q) { return UP ? last - q : first + q; };
  auto issue = [&](int q) {
    float* row_lds = ring[__builtin_amdgcn_readfirstlane(q % kDecR)];
    const float* row = px<float>(a.in, a.in_pitch, clampi(row_of(q), 0, a.nr - 1), 0);
    if constexpr (WIDE) {
      __builtin_amdgcn_global_load_lds(glb_addr(row + ic0 + 4 * lane), lds_addr(row_lds + kDecHL), 16u, 0, 0);
    } else {
#pragma unroll
      for (int m = 0; m < 4; m++)
        __builtin_amdgcn_global_load_lds(glb_addr(row + min(ic0 + 64 * m + lane, a.nc - 1)),
                                         lds_addr(row_lds + kDecHL + 64 * m), 4u, 0, 0);
    }
    if (lane < K) {
      __builtin_amdgcn_global_load_lds(glb_addr(row + hcl), lds_addr(row_lds + kDecHL - K), 4u, 0, 0);
      __builtin_amdgcn_global_load_lds(glb_addr(row + hcr), lds_addr(row_lds + kDecHL + kDecSW), 4u, 0, 0);
    }
  };
  auto store = [&](int o, float v0, float v1) {
    float* orow = px<float>(a.out, a.out_pitch, o, 0);
    if (WIDE && a.out_vec_ok) {
      const v2f_t w = {v0, v1};
      *reinterpret_cast<v2f_t*>(orow + oc) = w;
    } else {
      if (oc < a.onc) orow[oc] = v0;
      if (oc + 1 < a.onc) orow[oc + 1] = v1;
    }
  };
#pragma unroll
  for (int d = 0; d < kDecD; d++)
    if (d < nsteps) issue(d);

  float hist[P][2];
  float pend0 = 0.f, pend1 = 0.f;  // output row of the previous step, stored after this step's wait
  int pend_row = -1;
  for (int qb = 0; qb < nsteps; qb += P) {
    static_for<P>([&](auto tc) {
      constexpr int t = decltype(tc)::value;
      const int q = qb + t;
      if (q >= nsteps) return;
      if (q + kDecD - 1 >= nsteps) wait_vm<0>();
      else wait_vm<OPS * (kDecD - 1)>();
      if (pend_row >= 0) store(pend_row, pend0, pend1);
      pend_row = -1;
      // staged elements 4 lane .. 4 lane + 19; x[m] = staged element HL - K + 4 lane + m,
      // input column ic0 + 4 lane - K + m
      float4 ch[5];
      lds_read16_run<5>(lds_u32(ring[q % kDecR]) + 16u * (uint32_t)lane, ch);
      const float* x = reinterpret_cast<const float*>(ch) + (kDecHL - K);
      // the slot of step q + kDecD held step q - 1, read by the previous step
      if (q + kDecD < nsteps) issue(q + kDecD);
      // row pass at input columns ic0 + 4 lane and ic0 + 4 lane + 2
      float s0 = c[2 * K] * x[2 * K], s1 = c[2 * K] * x[2 * K + 2];
#pragma unroll
      for (int m = 2 * K - 1; m >= 0; m--) {
        s0 = fmaf(c[m], x[m], s0);
        s1 = fmaf(c[m], x[m + 2], s1);
      }
      hist[t][0] = s0;
      hist[t][1] = s1;
      // every other step from 2K on completes an output row o, whose tap m
      // (weight c[m]) is input row 2 o - K + m, added from m = 2K down to 0.
      // Going down, step q stages row 2 o + K and tap m was filtered at step
      // q - 2K + m (slot (t + m + 1) mod P); going up, step q stages row
      // 2 o - K and tap m was filtered at step q - m (slot (t + P - m) mod P).
      if (q >= 2 * K && ((q - 2 * K) & 1) == 0) {
        float v0, v1;
        if constexpr (!UP) {
          v0 = c[2 * K] * hist[t][0];
          v1 = c[2 * K] * hist[t][1];
#pragma unroll
          for (int m = 2 * K - 1; m >= 0; m--) {
            v0 = fmaf(c[m], hist[(t + m + 1) % P][0], v0);
            v1 = fmaf(c[m], hist[(t + m + 1) % P][1], v1);
          }
        } else {
          v0 = c[2 * K] * hist[(t + 1) % P][0];
          v1 = c[2 * K] * hist[(t + 1) % P][1];
#pragma unroll
          for (int m = 2 * K - 1; m >= 0; m--) {
            v0 = fmaf(c[m], hist[(t + P - m) % P][0], v0);
            v1 = fmaf(c[m], hist[(t + P - m) % P][1], v1);
          }
        }
        pend0 = v0;
        pend1 = v1;
        pend_row = UP ? o1 - 1 - (q - 2 * K) / 2 : o0 + (q - 2 * K) / 2;
      }
    });
  }
  if (pend_row >= 0) store(pend_row, pend0, pend1);
}

__global__ __launch_bounds__(64) void gauss_decimate_dma_kernel(dec_args a) {
  __shared__ __align__(16) float ring[kDecR * kDecRW];
  const int b = xcd_order(blockIdx.x, gridDim.x);
  const int tx = b % a.tiles_x, ty = b / a.tiles_x;
  const bool wide = a.vec_ok && (tx + 1) * kDecSW <= a.nc;
  if (ty & 1) {
    if (wide) dec_tile_dma<true, true>(a, tx, ty, ring);
    else dec_tile_dma<false, true>(a, tx, ty, ring);
  } else {
    if (wide) dec_tile_dma<true, false>(a, tx, ty, ring);
    else dec_tile_dma<false, false>(a, tx, ty, ring);
  }
}

static int gauss_decimate(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out) {
  if ((in->dtype != CUIMG_U8 && in->dtype != CUIMG_F32) || out->dtype != CUIMG_F32)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "pyr_down: Gaussian decimation reads U8 or F32 and writes F32");
  if (out->nrows != in->nrows / 2 || out->ncols != in->ncols / 2)
    return set_error(ctx, CUIMG_E_INVALID, "pyr_down: output is %dx%d, expected %dx%d", out->nrows, out->ncols,
                     in->nrows / 2, in->ncols / 2);
  dec_args a{};
  CUIMG_TRY(gaussian_coefs(1, kDecK, a.cf.c));
  a.in = in->begin;
  a.in_pitch = in->pitch;
  a.nr = in->nrows;
  a.nc = in->ncols;
  a.out = static_cast<float*>(out->begin);
  a.out_pitch = out->pitch;
  a.onr = out->nrows;
  a.onc = out->ncols;
  if (a.onr == 0 || a.onc == 0) return 0;
  if ((uint64_t)in->nrows * (uint64_t)in->pitch >= (1ull << 32))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "pyr_down: Gaussian decimation needs an input below 4 GiB");
  a.vec_ok = aligned16(in->begin, in->pitch);
  a.out_vec_ok = ((uintptr_t)out->begin % 8 == 0) && (out->pitch % 8 == 0);
  if (in->dtype == CUIMG_F32) {
    // streamed strips of 128 output columns
    a.tiles_x = ceil_div(a.onc, kDecSW / 2);
    a.tiles_y = ceil_div(a.onr, kDecSTH);
    const int blocks = a.tiles_x * a.tiles_y;
    return launch(ctx, "pyr_gauss_decimate",
                  [&] { gauss_decimate_dma_kernel<<<blocks, 64, 0, ctx->stream>>>(a); });
  }
  a.tiles_x = ceil_div(a.onc, kDecTW);
  a.tiles_y = ceil_div(a.onr, kDecTH);
  a.runs_y = ceil_div(a.tiles_y, kDecRun);
  const int blocks = a.tiles_x * a.runs_y;
  return launch(ctx, "pyr_gauss_decimate",
                [&] { gauss_decimate_kernel<uint8_t, false><<<blocks, 256, 0, ctx->stream>>>(a); });
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

int cuimg_hip_pyramid_build(cuimg_hip_ctx* ctx, const cuimg_img* frame, const cuimg_img* levels,
                            int nlevels) {
  return pyramid_build(ctx, frame, levels, nlevels);
}

int cuimg_hip_pyr_down(cuimg_hip_ctx* ctx, const cuimg_img* in, const cuimg_img* out, int mode) {
  if (!in || !out || !in->begin || !out->begin) return set_error(ctx, CUIMG_E_INVALID, "pyr_down: bad arguments");
  if (mode == CUIMG_PYR_GAUSS_DECIMATE) return gauss_decimate(ctx, in, out);
  if (mode != CUIMG_PYR_BOX2X2) return set_error(ctx, CUIMG_E_INVALID, "pyr_down: mode %d unknown", mode);
  // a two-level pyramid whose level 0 is the input itself
  // subsample writes the output domain only: neither border is touched
  cuimg_img lv[2] = {*in, *out};
  lv[0].border = 0;
  lv[1].border = 0;
  return pyramid_build(ctx, in, lv, 2);
}

}  // extern "C"
