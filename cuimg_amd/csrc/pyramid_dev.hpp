// Box pyramid for gfx950: the whole pyramid in one launch.
//
// Semantics: mipmap_kernel (cuimg/gpu/mipmap.h:25-50) applied level after
// level as tracker::subsample_input does (tracking2/tracker.hpp:63-69):
//   out(p) = (in(2p) + in(2p+(0,1)) + in(2p+(1,0)) + in(2p+(1,1))) / 4,
// integer sum and integer division for u8 (gl8u promotes to int), FP32 sum in
// that order then / 4 for float; level l+1 has floor(n_l / 2) rows and cols.
// Every level's border is then clamp-filled (cuimg/border.h:128-137).
//
// Design: a 256-thread workgroup owns a TW x TH tile of level 0 (at most 16 KiB,
// TH >= 2^(n-1)) and the nested TW/2 x TH/2, ... tiles of the coarser levels
// (2x2 blocks never cross tiles).  Full 16-byte chunks of the tile stream
// straight into LDS (global_load_lds, all of a thread's chunks in flight at
// once), ragged edges go through registers.  Level 0 is written back (unless
// the frame is level 0 itself) and each coarser level is built from the LDS copy
// of the previous one and written with its share of the clamped border.  HBM
// traffic is one read of level 0 plus one write of every produced level.  Wide,
// short tiles make each wave's loads whole row segments (one 1 KiB row per
// wave instruction for f32 256-wide tiles); images too large to stay in cache
// are written with non-temporal stores.
#pragma once

#include "common.hpp"

namespace cuimg {

constexpr int kPyrMaxLevels = 8;
constexpr int kPyrThreads = 256;          // workgroup size of every launch of pyramid_tile
constexpr int kPyrTileBytes = 16384;      // LDS of the level-0 tile

struct pyr_level {
  void* begin;
  int pitch, nr, nc, border;
};

struct pyr_args {
  const void* frame;
  int frame_pitch;
  int frame_vec_ok;   // frame rows 16-byte aligned
  int l0_vec_ok;      // level-0 rows 16-byte aligned
  int copy_level0;    // frame is not level 0 itself
  int n;
  int tiles_x;
  int tw_log2, th_log2; // level-0 tile: 2^tw_log2 columns x 2^th_log2 rows
  int stream_out;       // non-temporal stores
  pyr_level lv[kPyrMaxLevels];
};

// Host: the tile shape for an nr x nc level 0 of es-byte pixels and n levels
// (pyramid.hip).  Sets a.tw_log2, a.th_log2 and a.tiles_x; returns the number
// of tiles.
int pyr_tiling(pyr_args& a, int nr, int nc, int es);

template <typename T>
struct pyr_ops;
template <>
struct pyr_ops<uint8_t> {
  __device__ static uint8_t mean4(uint8_t a, uint8_t b, uint8_t c, uint8_t d) {
    return (uint8_t)((int(a) + int(b) + int(c) + int(d)) / 4);
  }
};
template <>
struct pyr_ops<float> {
  __device__ static float mean4(float a, float b, float c, float d) {
    float s = a + b;
    s = s + c;
    s = s + d;
    return s / 4;
  }
};

// Writes the border pixels of level L whose clamped source lies in the tile
// rows [r0,r1) x cols [c0,c1) (the tile's share of fill_border_clamp).
template <typename T>
__device__ inline void tile_border(const pyr_level& L, const T* tile, int tw, int r0, int r1, int c0, int c1) {
  const int b = L.border;
  if (b == 0) return;
  const int er0 = r0 == 0 ? -b : r0, er1 = r1 == L.nr ? L.nr + b : r1;
  const int ec0 = c0 == 0 ? -b : c0, ec1 = c1 == L.nc ? L.nc + b : c1;
  if (er0 == r0 && er1 == r1 && ec0 == c0 && ec1 == c1) return;
  const int w = ec1 - ec0, h = er1 - er0;
  for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
    const int r = er0 + i / w, c = ec0 + i % w;
    if (r >= 0 && r < L.nr && c >= 0 && c < L.nc) continue;
    const int rr = clampi(r, 0, L.nr - 1), cc = clampi(c, 0, L.nc - 1);
    *px<T>(L.begin, L.pitch, r, c) = tile[(rr - r0) * tw + (cc - c0)];
  }
}

template <typename T>
__device__ __forceinline__ void pyr_store(T* p, T v, bool nt) {
  if (nt) store_nt(p, v);
  else *p = v;
}
__device__ __forceinline__ void pyr_store16(void* p, uint4 v, bool nt) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u w = {v.x, v.y, v.z, v.w};
  if (nt) store_nt(p, __builtin_bit_cast(v4f_t, w));
  else *reinterpret_cast<v4u*>(p) = w;
}

// One level-0 tile (and its nested coarser tiles) per kPyrThreads workgroup.
template <typename T>
__device__ inline void pyramid_tile(const pyr_args& a, int tile) {
  __shared__ __align__(16) T t0[kPyrTileBytes / sizeof(T)];
  __shared__ __align__(16) T t1[kPyrTileBytes / 4 / sizeof(T)];
  T* cur = t0;
  T* nxt = t1;

  const int twl = a.tw_log2, thl = a.th_log2;
  const int TW = 1 << twl, TH = 1 << thl;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const pyr_level& L0 = a.lv[0];
  const int r0 = ty * TH, c0 = tx * TW;
  const int r1 = min(r0 + TH, L0.nr), c1 = min(c0 + TW, L0.nc);
  const bool nt = a.stream_out;

  // ---- level-0 tile: chunk i covers tile row i / CPR, columns (i % CPR) * CH ..
  constexpr int CH = 16 / sizeof(T);        // elements per 16-byte chunk
  constexpr int CH_LOG2 = sizeof(T) == 1 ? 4 : 2;
  static_assert(sizeof(T) == 1 || sizeof(T) == 4, "u8 or f32 pixels");
  constexpr int NPT = kPyrTileBytes / 16 / kPyrThreads;  // chunks per thread, at most
  const int cprl = twl - CH_LOG2;
  const int nchunks = 1 << (twl + thl - CH_LOG2);
  if (a.frame_vec_ok && c1 - c0 == TW) {
    // whole chunks: the tile is dense in LDS, so the 64 chunks of one wave
    // instruction land at consecutive 16-byte slots from the wave's first one
#pragma unroll
    for (int k = 0; k < NPT; k++) {
      const int i = threadIdx.x + kPyrThreads * k;
      if (i < nchunks && r0 + (i >> cprl) < r1) {
        const int wi = __builtin_amdgcn_readfirstlane(i & ~63);
        const T* g = px<T>(a.frame, a.frame_pitch, r0 + (i >> cprl), c0 + ((i & ((1 << cprl) - 1)) << CH_LOG2));
        __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void*)g,
                                         (__attribute__((address_space(3))) void*)(t0 + (wi << CH_LOG2)), 16u, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0 | (7 << 4) | (15 << 8));  // vmcnt(0)
    if (a.copy_level0) {
      const bool l0_vec = a.l0_vec_ok;
#pragma unroll
      for (int k = 0; k < NPT; k++) {
        const int i = threadIdx.x + kPyrThreads * k;
        if (i < nchunks && r0 + (i >> cprl) < r1) {
          // own chunk: written by this lane's DMA, complete after the wait
          const uint4 v = *reinterpret_cast<const uint4*>(t0 + (i << CH_LOG2));
          T* dst = px<T>(L0.begin, L0.pitch, r0 + (i >> cprl), c0 + ((i & ((1 << cprl) - 1)) << CH_LOG2));
          if (l0_vec) {
            pyr_store16(dst, v, nt);
          } else {
            // level 0 with an odd border: element stores
            const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
            for (int m = 0; m < CH; m++) pyr_store(dst + m, e[m], nt);
          }
        }
      }
    }
  } else {
    for (int i = threadIdx.x; i < nchunks; i += kPyrThreads) {
      const int lr = i >> cprl, lc = (i & ((1 << cprl) - 1)) << CH_LOG2;
      const int r = r0 + lr, c = c0 + lc;
      if (r >= r1 || c >= c1) continue;
      for (int k = 0; k < CH && c + k < c1; k++) {
        const T v = *px<T>(a.frame, a.frame_pitch, r, c + k);
        t0[(lr << twl) + lc + k] = v;
        if (a.copy_level0) pyr_store(px<T>(L0.begin, L0.pitch, r, c + k), v, nt);
      }
    }
  }
  __syncthreads();
  tile_border<T>(L0, t0, TW, r0, r1, c0, c1);

  // ---- coarser levels from the LDS copy of the previous level
  int wl = twl, hl = thl;
  for (int l = 1; l < a.n; l++) {
    const pyr_level& L = a.lv[l];
    const int nwl = wl - 1, nhl = hl - 1;
    const int ntw = 1 << nwl;
    const int lr0 = ty << nhl, lc0 = tx << nwl;
    const int lr1 = min(lr0 + (1 << nhl), L.nr), lc1 = min(lc0 + ntw, L.nc);
    for (int i = threadIdx.x; i < (1 << (nwl + nhl)); i += kPyrThreads) {
      const int r = i >> nwl, c = i & (ntw - 1);
      if (lr0 + r >= lr1 || lc0 + c >= lc1) continue;
      const T* s = &cur[((2 * r) << wl) + 2 * c];
      T v = pyr_ops<T>::mean4(s[0], s[1], s[1 << wl], s[(1 << wl) + 1]);
      nxt[(r << nwl) + c] = v;
      pyr_store(px<T>(L.begin, L.pitch, lr0 + r, lc0 + c), v, nt);
    }
    __syncthreads();
    if (lr0 < lr1 && lc0 < lc1) tile_border<T>(L, nxt, ntw, lr0, lr1, lc0, lc1);
    __syncthreads();
    T* t = cur;
    cur = nxt;
    nxt = t;
    wl = nwl;
    hl = nhl;
  }
}

}  // namespace cuimg
