// Box pyramid for gfx950: the whole pyramid in one launch.
//
// Semantics: mipmap_kernel (cuimg/gpu/mipmap.h:25-50) applied level after
// level as tracker::subsample_input does (tracking2/tracker.hpp:63-69):
//   out(p) = (in(2p) + in(2p+(0,1)) + in(2p+(1,0)) + in(2p+(1,1))) / 4,
// integer sum and integer division for u8 (gl8u promotes to int), FP32 sum in
// that order then / 4 for float; level l+1 has floor(n_l / 2) rows and cols.
// Every level's border is then clamp-filled (cuimg/border.h:128-137).
//
// Design: a 256-thread workgroup owns a 64x64 tile of level 0 and the nested
// 32x32, 16x16, ... tiles of the coarser levels (2x2 blocks never cross tiles).
// It reads the tile once with 16-byte loads into LDS, writes level 0 (unless
// level 0 is the source itself), and builds each coarser level from the LDS copy
// of the previous one, writing it with its share of the clamped border.  HBM
// traffic is one read of level 0 plus one write of every produced level.
#pragma once

#include "common.hpp"

namespace cuimg {

constexpr int kPyrMaxLevels = 8;
constexpr int kTile = 64;

struct pyr_level {
  void* begin;
  int pitch, nr, nc, border;
};

struct pyr_args {
  const void* frame;
  int frame_pitch;
  int frame_vec_ok;   // frame and level 0 rows 16-byte aligned
  int copy_level0;    // frame is not level 0 itself
  int n;
  int tiles_x;
  pyr_level lv[kPyrMaxLevels];
};

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

// One 64x64 level-0 tile (and its nested coarser tiles) per 256-thread
// workgroup.
template <typename T>
__device__ inline void pyramid_tile(const pyr_args& a, int tile) {
  __shared__ T t0[kTile * kTile];
  __shared__ T t1[(kTile / 2) * (kTile / 2)];
  T* cur = t0;
  T* nxt = t1;

  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const pyr_level& L0 = a.lv[0];
  const int r0 = ty * kTile, c0 = tx * kTile;
  const int r1 = min(r0 + kTile, L0.nr), c1 = min(c0 + kTile, L0.nc);

  // ---- level 0 tile: load (and copy when the frame is a separate image)
  constexpr int CH = 16 / sizeof(T);           // elements per 16-byte chunk
  constexpr int CPR = kTile / CH;              // chunks per tile row
  if (a.frame_vec_ok && c1 - c0 == kTile) {
    for (int i = threadIdx.x; i < kTile * CPR; i += blockDim.x) {
      const int lr = i / CPR, lc = (i % CPR) * CH;
      if (r0 + lr >= r1) continue;
      uint4 v = *reinterpret_cast<const uint4*>(px<T>(a.frame, a.frame_pitch, r0 + lr, c0 + lc));
      *reinterpret_cast<uint4*>(&t0[lr * kTile + lc]) = v;
      if (a.copy_level0) *reinterpret_cast<uint4*>(px<T>(L0.begin, L0.pitch, r0 + lr, c0 + lc)) = v;
    }
  } else {
    for (int i = threadIdx.x; i < kTile * kTile; i += blockDim.x) {
      const int lr = i / kTile, lc = i % kTile;
      if (r0 + lr >= r1 || c0 + lc >= c1) continue;
      T v = *px<T>(a.frame, a.frame_pitch, r0 + lr, c0 + lc);
      t0[lr * kTile + lc] = v;
      if (a.copy_level0) *px<T>(L0.begin, L0.pitch, r0 + lr, c0 + lc) = v;
    }
  }
  __syncthreads();
  tile_border<T>(L0, t0, kTile, r0, r1, c0, c1);

  // ---- coarser levels from the LDS copy of the previous level
  int tw = kTile;
  for (int l = 1; l < a.n; l++) {
    const pyr_level& L = a.lv[l];
    const int ntw = tw / 2;
    const int lr0 = ty * ntw, lc0 = tx * ntw;
    const int lr1 = min(lr0 + ntw, L.nr), lc1 = min(lc0 + ntw, L.nc);
    for (int i = threadIdx.x; i < ntw * ntw; i += blockDim.x) {
      const int r = i / ntw, c = i % ntw;
      if (lr0 + r >= lr1 || lc0 + c >= lc1) continue;
      const T* s = &cur[(2 * r) * tw + 2 * c];
      T v = pyr_ops<T>::mean4(s[0], s[1], s[tw], s[tw + 1]);
      nxt[r * ntw + c] = v;
      *px<T>(L.begin, L.pitch, lr0 + r, lc0 + c) = v;
    }
    __syncthreads();
    if (lr0 < lr1 && lc0 < lc1) tile_border<T>(L, nxt, ntw, lr0, lr1, lc0, lc1);
    __syncthreads();
    T* t = cur;
    cur = nxt;
    nxt = t;
    tw = ntw;
  }
}

}  // namespace cuimg
