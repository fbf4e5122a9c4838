// FAST-like saliency for gfx950 (integer, bit-exact with the reference CPU path).
//
// Reference: fast::compute_saliency (tracking2/fast_detector.hpp:26-157) run by
// compute_saliency_kernel wherever the detector mask is non-zero (:267-313).
// For each pixel p with value vp and the 16 points of the radius-3 ring
// (cuimg/neighb2d_data.h:184-189), a ring point is "bright" if v - vp > th and
// "dark" if vp - v > th.  The reference scans the ring keeping the longest run
// of equal-sign beyond-threshold points, continuing the last run through index
// 0; the score is 0 when that run is shorter than n, else
// max(sum of bright distances, sum of dark distances) truncated to 8 bits.
// Its early test on points {0,4,8,12,0,4} (n >= 9) only rejects pixels whose
// longest run is below n, so it does not change any score and is not repeated.
//
// Here the run test is branch-free: with the 16-bit masks B (bright) and D
// (dark), a circular run of length >= n exists iff the AND of the mask with
// its n-1 successive rotations is non-zero, computed in log2(n) + 1 shift/AND
// steps on the mask repeated over 32 bits.  A 256-thread workgroup stages a
// (16+6) x (64+6) byte tile in LDS (one global read per input byte) and each
// thread scores 4 pixels of one column.
#pragma once

#include "common.hpp"

namespace cuimg {

// ring offsets (row, col), CPU table order
static __constant__ int8_t kRingR[16] = {-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3};
static __constant__ int8_t kRingC[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};

constexpr int kFTW = 64, kFTH = 16, kFH = 3;

// A circular run of n (1..16) set bits in the 16-bit mask m.  On m repeated
// over 32 bits, bit j of r marks a run of `len` set bits starting at j; the
// run doubles per step, and a last step of n - len <= len completes it.  Bits
// 0..15 of the result cover every start of the circular mask.
__device__ __forceinline__ bool has_run(unsigned m, int n) {
  unsigned r = m | (m << 16);
  int len = 1;
  while (2 * len <= n) {
    r &= r >> len;
    len *= 2;
  }
  if (n > len) r &= r >> (n - len);
  return (r & 0xFFFFu) != 0;
}

__device__ __forceinline__ uint8_t fast_score_lds(const uint8_t* t, int stride, int n, int th) {
  const int vp = t[0];
  unsigned B = 0, D = 0, sb = 0, sd = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int v = t[kRingR[i] * stride + kRingC[i]];
    const int d = v - vp;
    if (d > th) {
      B |= 1u << i;
      sb += (unsigned)d;
    }
    if (-d > th) {
      D |= 1u << i;
      sd += (unsigned)(-d);
    }
  }
  if (n > 16) return 0;
  if (!has_run(B, n) && !has_run(D, n)) return 0;
  return (uint8_t)(sb > sd ? sb : sd);
}

// One (16 x 64)-pixel tile per 256-thread workgroup.  A pixel is scored where
// its mask byte is at least mask_min (a 0 / 255 detector mask: 1; the bc2s
// strategy's contrast, saturated to 8 bits, as its mask: 5).
template <bool CLAMP>
__device__ inline void fast_tile_body(const uint8_t* in, int in_pitch, const uint8_t* mask, int mask_pitch,
                                      uint8_t* out, int out_pitch, int nr, int nc, int tile, int tiles_x, int n,
                                      int th, int mask_min = 1) {
  constexpr int SW = kFTW + 2 * kFH, SH = kFTH + 2 * kFH;
  __shared__ uint8_t t[SH * SW];
  const int r0 = (tile / tiles_x) * kFTH, c0 = (tile % tiles_x) * kFTW;
  const int lc = threadIdx.x % kFTW, lr0 = threadIdx.x / kFTW;
  const int c = c0 + lc;
  // the mask bytes of this thread's pixels, loaded with the tile (one memory
  // round trip for both)
  uint8_t mk[kFTH / 4];
#pragma unroll
  for (int k = 0; k < kFTH / 4; k++) {
    const int r = r0 + lr0 + 4 * k;
    mk[k] = (!mask || r >= nr || c >= nc) ? 255 : *px<uint8_t>(mask, mask_pitch, r, c);
  }
  // the tile with its halo: every load of the thread issued before the first
  // LDS write (one memory round trip)
  constexpr int NL = (SH * SW + 255) / 256;
  uint8_t v[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int i = k * 256 + threadIdx.x;
    int rr = r0 - kFH + i / SW, cc = c0 - kFH + i % SW;
    if (CLAMP) {
      rr = clampi(rr, 0, nr - 1);
      cc = clampi(cc, 0, nc - 1);
    } else {
      // the image border (>= 3, clamp-filled) covers the ring; rows/cols past
      // a partial tile are clamped so they stay inside the allocation
      rr = min(rr, nr + kFH - 1);
      cc = min(cc, nc + kFH - 1);
    }
    if (i < SH * SW) v[k] = *px<uint8_t>(in, in_pitch, rr, cc);
  }
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int i = k * 256 + threadIdx.x;
    if (i < SH * SW) t[i] = v[k];
  }
  __syncthreads();
  if (c >= nc) return;
#pragma unroll
  for (int k = 0; k < kFTH / 4; k++) {
    const int lr = lr0 + 4 * k, r = r0 + lr;
    if (r >= nr) break;
    uint8_t s = 0;
    if (mk[k] >= mask_min) s = fast_score_lds(&t[(lr + kFH) * SW + lc + kFH], SW, n, th);
    *px<uint8_t>(out, out_pitch, r, c) = s;
  }
}

}  // namespace cuimg
