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
// steps on the mask repeated over 32 bits.  fast_tile_body (any threshold,
// any image) stages a (16+6) x (64+6) byte tile in LDS (one global read per
// input byte) and each thread scores 4 pixels of one column; fast_tile4_body
// (th >= 0, the tracker's images) scores 4 pixels of one row per thread in
// packed 16-bit arithmetic.
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

// ------------------------------------------------- four pixels per thread
// The same scores for 0 <= th, four horizontally adjacent pixels per thread,
// two at a time as the 16-bit halves of one register (v_pk_* integer ops).
// With hi = vp + th and lo = max(vp - th, 0):
//   t_b = max(v - hi, 0) > 0  iff  v - vp > th   (bright), t_b = d - th,
//   t_d = max(lo - v, 0) > 0  iff  vp - v > th   (dark),   t_d = -d - th,
// so the bright sum is sum(t_b) + th * |B| and the dark sum sum(t_d) + th * |D|
// (th > 255 marks nothing, as min(th, 255) does).  Mask bit i is
// min(t, 1) << i (no two points share a bit, so the bits are added); the run test rotates each 16-bit half (rotr by len ANDed in,
// len doubling, as has_run).  A ring row of the quad is three LDS words, and
// v_perm picks each ring point's two pixel pairs out of two of them.
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2_t pk(uint32_t x) { return __builtin_bit_cast(u16x2_t, x); }
__device__ __forceinline__ uint32_t upk(u16x2_t x) { return __builtin_bit_cast(uint32_t, x); }

// Per 16-bit half: rotate right by s (1 <= s <= 15).
__device__ __forceinline__ uint32_t rotr16x2(uint32_t x, int s) {
  const u16x2_t v = pk(x);
  const u16x2_t a = {(unsigned short)s, (unsigned short)s};
  const u16x2_t b = {(unsigned short)(16 - s), (unsigned short)(16 - s)};
  return upk((v >> a) | (v << b));
}

// has_run on both halves: the result's half is non-zero iff that half's mask
// holds a circular run of n set bits.
__device__ __forceinline__ uint32_t has_run2(uint32_t m, int n) {
  uint32_t r = m;
  int len = 1;
  while (2 * len <= n) {
    r &= rotr16x2(r, len);
    len *= 2;
  }
  if (n > len) r &= rotr16x2(r, n - len);
  return r;
}

constexpr int kF4SW = kFTW + 8;  // LDS row: global columns c0 - 4 .. c0 + 67, 4-byte aligned

// One (16 x 64)-pixel tile per 256-thread workgroup, th >= 0.  The image rows
// have a clamp-filled border of at least 3 and at least 4 bytes before column
// 0, and the rows of in, mask and out are 4-byte aligned at column 0; the
// tile reads columns up to c0 + 67 of rows up to nr + 2 (bytes past column
// nc + 2 only feed pixels outside the domain).
__device__ inline void fast_tile4_body(const uint8_t* in, int in_pitch, const uint8_t* mask, int mask_pitch,
                                       uint8_t* out, int out_pitch, int nr, int nc, int tile, int tiles_x, int n,
                                       int th, int mask_min) {
  constexpr int SH = kFTH + 2 * kFH, NW = kF4SW / 4;
  __shared__ uint32_t t[SH * NW];
  const int r0 = (tile / tiles_x) * kFTH, c0 = (tile % tiles_x) * kFTW;
  const int lr = threadIdx.x / (kFTW / 4), q = threadIdx.x % (kFTW / 4);
  const int r = r0 + lr, c = c0 + 4 * q;
  uint32_t mk = 0xFFFFFFFFu;
  if (mask && r < nr) {
    const uint8_t* mp = px<uint8_t>(mask, mask_pitch, r, c);
    if (c + 4 <= nc) {
      mk = *reinterpret_cast<const uint32_t*>(mp);
    } else {
      mk = 0;
      for (int j = 0; j < nc - c; j++) mk |= (uint32_t)mp[j] << (8 * j);
    }
  }
  constexpr int NL = (SH * NW + 255) / 256;
  uint32_t v[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int i = k * 256 + threadIdx.x;
    if (i < SH * NW) {
      const int rr = min(r0 - kFH + i / NW, nr + kFH - 1);
      v[k] = *reinterpret_cast<const uint32_t*>(px<uint8_t>(in, in_pitch, rr, c0 - 4 + 4 * (i % NW)));
    }
  }
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int i = k * 256 + threadIdx.x;
    if (i < SH * NW) t[i] = v[k];
  }
  __syncthreads();
  if (r >= nr || c >= nc) return;
  uint32_t res = 0;
  if (n <= 16) {
    constexpr int8_t RR[16] = {-3, -3, -2, -1, 0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3};
    constexpr int8_t RC[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
    uint32_t w[7][3];
#pragma unroll
    for (int d = 0; d < 7; d++)
#pragma unroll
      for (int k = 0; k < 3; k++) w[d][k] = t[(lr + d) * NW + q + k];
    // pixel j of the quad is LDS byte 4 + j of its row's three words
    const uint32_t vp0 = __builtin_amdgcn_perm(w[3][2], w[3][1], 0x0c010c00u);
    const uint32_t vp1 = __builtin_amdgcn_perm(w[3][2], w[3][1], 0x0c030c02u);
    const unsigned short th16 = (unsigned short)min(th, 255);
    const u16x2_t thv = {th16, th16};
    // (1, 1) in a register the optimizer cannot see through: min(t, 1) then
    // stays one v_pk_min_u16 instead of a compare and a select per half
    uint32_t one_raw = 0x00010001u;
    asm volatile("" : "+v"(one_raw));
    const u16x2_t one = pk(one_raw);
    const u16x2_t hi0 = pk(vp0) + thv, hi1 = pk(vp1) + thv;
    const u16x2_t lo0 = __builtin_elementwise_sub_sat(pk(vp0), thv), lo1 = __builtin_elementwise_sub_sat(pk(vp1), thv);
    u16x2_t sb0 = {0, 0}, sb1 = {0, 0}, sd0 = {0, 0}, sd1 = {0, 0};
    uint32_t B0 = 0, B1 = 0, D0 = 0, D1 = 0;
    static_for<16>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      constexpr int o = 4 + RC[i];         // LDS byte of pixel 0's ring point
      constexpr int wl = o >> 2, O = o & 3;  // pairs from words wl, wl + 1
      constexpr uint32_t s0 = 0x0c000c00u | (uint32_t)O | ((uint32_t)(O + 1) << 16);
      constexpr uint32_t s1 = 0x0c000c00u | (uint32_t)(O + 2) | ((uint32_t)(O + 3) << 16);
      const uint32_t lw = w[3 + RR[i]][wl], hw = w[3 + RR[i]][wl + 1];
      const u16x2_t x0 = pk(__builtin_amdgcn_perm(hw, lw, s0)), x1 = pk(__builtin_amdgcn_perm(hw, lw, s1));
      const u16x2_t tb0 = __builtin_elementwise_sub_sat(x0, hi0), tb1 = __builtin_elementwise_sub_sat(x1, hi1);
      const u16x2_t td0 = __builtin_elementwise_sub_sat(lo0, x0), td1 = __builtin_elementwise_sub_sat(lo1, x1);
      sb0 += tb0;
      sb1 += tb1;
      sd0 += td0;
      sd1 += td1;
      // disjoint bits: the add is the OR (one v_lshl_add per mask)
      B0 += upk(__builtin_elementwise_min(tb0, one)) << i;
      B1 += upk(__builtin_elementwise_min(tb1, one)) << i;
      D0 += upk(__builtin_elementwise_min(td0, one)) << i;
      D1 += upk(__builtin_elementwise_min(td1, one)) << i;
    });
    const uint32_t run0 = has_run2(B0, n) | has_run2(D0, n);
    const uint32_t run1 = has_run2(B1, n) | has_run2(D1, n);
    const uint32_t B[2] = {B0, B1}, D[2] = {D0, D1}, run[2] = {run0, run1};
    const uint32_t sb[2] = {upk(sb0), upk(sb1)}, sd[2] = {upk(sd0), upk(sd1)};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int p = j >> 1, h = 16 * (j & 1);
      const uint32_t b = (sb[p] >> h & 0xFFFFu) + (uint32_t)th16 * (uint32_t)__popc(B[p] >> h & 0xFFFFu);
      const uint32_t dd = (sd[p] >> h & 0xFFFFu) + (uint32_t)th16 * (uint32_t)__popc(D[p] >> h & 0xFFFFu);
      const uint32_t s = (run[p] >> h & 0xFFFFu) ? ((b > dd ? b : dd) & 0xFFu) : 0u;
      const uint32_t m = mk >> (8 * j) & 0xFFu;
      res |= (m >= (uint32_t)mask_min ? s : 0u) << (8 * j);
    }
  }
  uint8_t* op = px<uint8_t>(out, out_pitch, r, c);
  if (c + 4 <= nc) {
    *reinterpret_cast<uint32_t*>(op) = res;
  } else {
    for (int j = 0; j < nc - c; j++) op[j] = (uint8_t)(res >> (8 * j));
  }
}

}  // namespace cuimg
