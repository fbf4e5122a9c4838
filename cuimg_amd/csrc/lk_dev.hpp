// Pyramid Lucas-Kanade particle matcher for gfx950.
//
// Reference: pyrlk_match_particles_kernel (tracking2/pyrlk_tracker.hpp:676-769)
// with iterative_lk_match_5x5::match (:165-335), the bilinear sampler `li`
// (:130-146), the Sobel/8 gradient of the previous frame (:909-938) and the
// recursive coarse-to-fine prediction (predictions.h:60-74,
// pyrlk_tracker.hpp:1025-1063, transformations.h:77-80).
//
// Design: a particle is matched by a group of GL lanes: 16 lanes (four
// particles per wave, one DPP row each) for windows up to 7x7, a whole wave for
// larger windows.  Lane l of a group owns window pixels l, l+GL, ... (raster
// order, NS per lane).  The per-particle scalar work (the 2x2 solve, the
// update, the tests) then runs once for four particles, and a window sum is a
// 4-step DPP reduction inside the row.  Groups of one wave diverge only as
// whole groups.  A level solve is latency-bound, so it is built around few
// memory round trips:
//  * the previous frame's (ws+5)^2 patch around the particle and the current
//    frame's (ws+1+2*kMB)^2 patch around the prediction are loaded into LDS in
//    one batch of clamped reads;
//  * the Sobel/8 gradient of the previous frame is computed for the (ws+3)^2
//    inner pixels of its patch in LDS (zero on the image's outer ring), so no
//    gradient image exists; template samples and gradients come from LDS;
//  * Gauss-Newton steps sample the current-frame patch; it is reloaded only when
//    the estimate leaves its kMB-pixel margin;
//  * window sums are group reductions (DPP within rows, and for 64-lane groups
//    two cross-row broadcasts), bit-identical in every lane of the group; the
//    2x2 solve runs redundantly per lane in FP32 registers, so all control
//    flow is uniform over a group.
// Sample rows/cols satisfy |tz(x + r) - (tz(x) + r)| <= 1 for float x and
// integer r, which sizes the patches: every corner a lane reads lies inside.
// No MFMA: nothing here is a dense contraction.  A particle at scale s first
// re-solves its synthetic ancestors at scales L-1 .. s+1, as the reference's
// recursion does, each solve feeding the next one's prediction.
//
// Numerics: the arithmetic is fixed operation by operation (contraction of a
// product into an add is off in this file, every FMA is written out):
//  * bilinear weights a0 = q - trunc(q), 1 - a0, ..., w00 = (1-a0)(1-a1), ...
//    each rounded; a sample is w00 v00, then fma(w10, v10, .), fma(w01, v01, .),
//    fma(w11, v11, .);
//  * Sobel sums, the 2x2 solve, the tests and the updates: each operation
//    rounded, in the reference's order;
//  * window sums: lane l of a group accumulates its pixels l, l + GL, ... in
//    ascending order (products with fma), then the group's lanes are added
//    pairwise by the DPP reductions below.
// The oracle's "GPU arithmetic" mode performs exactly these operations, so the
// two agree bit for bit.
#pragma once

#include "tracker_state.hpp"

#include <type_traits>

#pragma clang fp contract(off)

namespace cuimg {

// Single rounded operations (written here, under contraction off: the HIP
// header's __fmul_rn / __fadd_rn are plain operators compiled with
// contraction on, so a product of theirs can still be fused into an add).
__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }
__device__ __forceinline__ float sub_rn(float a, float b) { return a - b; }

// In-row permutations only (every source lane exists), so bound_ctrl changes no
// value; with it set the compiler folds the move into the consuming add.
template <int CTRL>
__device__ __forceinline__ float mov_dpp(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float upd_dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xF, false));
}

// Sum over the 64 lanes as one wave-uniform (scalar) value: xor-1 and xor-2
// quad permutes, half-row and row mirrors give every lane its row's sum, then
// row_bcast15 / row_bcast31 fold rows (1 += 0, 3 += 2, then 3 += 1) and lane
// 63 holds ((r3 + r2) + (r1 + r0)).  No LDS traffic.
__device__ __forceinline__ float wave_sum(float x) {
  x += mov_dpp<0xB1>(x);          // quad_perm [1,0,3,2]
  x += mov_dpp<0x4E>(x);          // quad_perm [2,3,0,1]
  x += mov_dpp<0x141>(x);         // row_half_mirror
  x += mov_dpp<0x140>(x);         // row_mirror
  x += upd_dpp<0x142, 0xA>(x);    // row_bcast:15 into rows 1, 3
  x += upd_dpp<0x143, 0xC>(x);    // row_bcast:31 into rows 2, 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Sum over a group of GL lanes (16: one DPP row; 64: the wave), the same bits
// in every lane of the group.
template <int GL>
__device__ __forceinline__ float group_sum(float x) {
  if constexpr (GL == 64) {
    return wave_sum(x);
  } else {
    static_assert(GL == 16, "16- or 64-lane groups");
    x += mov_dpp<0xB1>(x);   // quad_perm [1,0,3,2]
    x += mov_dpp<0x4E>(x);   // quad_perm [2,3,0,1]
    x += mov_dpp<0x141>(x);  // row_half_mirror
    x += mov_dpp<0x140>(x);  // row_mirror
    return x;
  }
}

// The bits of a wave ballot that belong to this lane's group.
template <int GL>
__device__ __forceinline__ uint64_t group_ballot(bool v) {
  const uint64_t b = __ballot(v);
  if constexpr (GL == 64) {
    return b;
  } else {
    const int sh = (int)(threadIdx.x & 63) & ~(GL - 1);
    return (b >> sh) & ((1ull << GL) - 1);
  }
}

__device__ __forceinline__ int tz(float v) { return (int)v; }


__device__ __forceinline__ float norml2(float a, float b) {
  return sqrtf(add_rn(mul_rn(a, a), mul_rn(b, b)));
}

// k / q for 0 <= k < 4096, q <= 64: (k + 0.5) / q is at least 0.5/q away from
// an integer, far more than the float error of the product.
__device__ __forceinline__ int divq(int k, float invq) { return (int)(((float)k + 0.5f) * invq); }


// Bilinear weights of li: base = coordinates truncated toward zero.
struct bw {
  int xr, xc;
  float w00, w10, w01, w11;
};

__device__ __forceinline__ bw weights(float qr, float qc) {
  bw w;
  w.xr = tz(qr);
  w.xc = tz(qc);
  const float a0 = sub_rn(qr, (float)w.xr), a1 = sub_rn(qc, (float)w.xc);
  const float b0 = sub_rn(1.f, a0), b1 = sub_rn(1.f, a1);
  w.w00 = mul_rn(b0, b1);
  w.w10 = mul_rn(a0, b1);
  w.w01 = mul_rn(b0, a1);
  w.w11 = mul_rn(a0, a1);
  return w;
}

// li's left-to-right sum w00 v00 + w10 v10 + w01 v01 + w11 v11, the last
// three terms fused.
__device__ __forceinline__ float blend(const bw& w, float v00, float v10, float v01, float v11) {
  float s = mul_rn(w.w00, v00);
  s = fmaf(w.w10, v10, s);
  s = fmaf(w.w01, v01, s);
  s = fmaf(w.w11, v11, s);
  return s;
}

struct match_out {
  float2 tr;
  float dist;
};

constexpr int kMB = 3;  // motion margin of the current-frame patch

// Rows of a patch of q rows with stride ps loaded by gl lanes (whole passes).
constexpr int patch_rows(int q, int ps, int gl) {
  return gl % ps == 0 ? (q + gl / ps - 1) / (gl / ps) * (gl / ps) : q;
}

// Waves per matcher workgroup.
constexpr int kLkWaves = 4;
// Matcher workgroups per job of a frame step: the grid cannot follow the
// device-side particle count, and every workgroup without particles still
// holds its LDS while it reads the count, so the grid stops well short of the
// capacity.
constexpr int kLkMaxWgs = 1024;
constexpr int kLkThreads = 64 * kLkWaves;

// Window class NPL (1, 2, 4: windows up to 7x7, 11x11, 15x15) and its group
// layout: GL lanes per particle, NS window pixels per lane.
template <int NPL>
struct lk_cfg {
  static constexpr int WSMAX = NPL == 1 ? 7 : (NPL == 2 ? 11 : 15);
  static constexpr int GL = NPL == 1 ? 16 : 64;
  static constexpr int NS = (WSMAX * WSMAX + GL - 1) / GL;
  static constexpr int GPW = 64 / GL;  // particles per wave in flight
};

// LDS of one group, sized for the largest window of its class.  The
// current-frame patch has a row stride of PW (16 or 32 floats) so a group
// loads GL/PW whole rows per pass with constant offsets; the previous-frame
// patch of a 16-lane group is stored densely (stride QA, one row per pass).
template <int NPL, bool SEL>
struct lk_lds {
  using C = lk_cfg<NPL>;
  static constexpr int WSMAX = C::WSMAX;
  static constexpr int QA = WSMAX + 5;            // previous-frame patch side
  static constexpr int QG = WSMAX + 3;            // gradient patch side
  static constexpr int QB = WSMAX + 1 + 2 * kMB;  // current-frame patch side
  static constexpr int PW = QB <= 16 ? 16 : 32;   // current-frame patch row stride (>= QB)
  static constexpr int PWA = C::GL == 16 ? QA : PW;  // previous-frame patch row stride (>= QA)
  static constexpr int GPASS = (QG * QG + C::GL - 1) / C::GL;
  static constexpr int FPASS = ((WSMAX + 1) * (WSMAX + 1) + C::GL - 1) / C::GL;
  float2 G[QG * QG];
  float A[patch_rows(QA, PWA, C::GL) * PWA];
  float B[patch_rows(QB, PW, C::GL) * PW];
  // template samples: the error term reads index (r+h)*h + (c+h), h = ws/2,
  // at most (ws-1)*h + ws-1, so only that prefix is kept
  static constexpr int TPL = (WSMAX - 1) * (WSMAX / 2) + WSMAX;
  float tpl[TPL];
  float mag[SEL ? C::GL * C::NS : 1];  // |gradient| ranks, top_n selection only
};

// dst[i*PS + j] = img[clamp(r0 + i), clamp(c0 + j)] for i, j < q <= MAXQ:
// indexing the patch with an unclamped coordinate reproduces the reference's
// clamped read.  With PS dividing GL, group lane l loads column l % PS of rows
// l / PS + m * (GL / PS); otherwise lane l loads column l of row m.  All of a
// lane's loads are issued before its first LDS write, so a patch costs one
// memory round trip.
template <typename TI, int PS, int GL, int MAXQ>
__device__ __forceinline__ void load_patch(float* dst, int q, const void* img, int pitch, int nr, int nc, int r0,
                                           int c0, int lane) {
  constexpr bool WHOLE = GL % PS == 0;
  constexpr int RPP = WHOLE ? GL / PS : 1;
  constexpr int MP = (MAXQ + RPP - 1) / RPP;
  const int col = WHOLE ? lane % PS : lane, row = WHOLE ? lane / PS : 0;
  const bool on = col < q && col < PS;
  // 32-bit byte offsets from the (wave-uniform) image base: one 24-bit
  // multiply-add per load (images are below 2^24 rows and 2^24 pitch bytes)
  const uint32_t cb = (uint32_t)clampi(c0 + col, 0, nc - 1) * (uint32_t)sizeof(TI);
  const char* base = static_cast<const char*>(img);
  float v[MP];
#pragma unroll
  for (int m = 0; m < MP; m++) {
    const int i = row + m * RPP;
    if (on && i < q) {
      const uint32_t off = __umul24((uint32_t)clampi(r0 + i, 0, nr - 1), (uint32_t)pitch) + cb;
      v[m] = (float)*reinterpret_cast<const TI*>(base + off);
    }
  }
#pragma unroll
  for (int m = 0; m < MP; m++) {
    const int i = row + m * RPP;
    if (on && i < q) dst[i * PS + col] = v[m];
  }
}

// load_patch for a square FP32 patch of Q = PS whole rows through a raw
// buffer view: row offsets are one 24-bit multiply-add, no 64-bit address
// arithmetic, and every lane loads and stores unconditionally.
template <int PS, int GL>
__device__ __forceinline__ void load_patch_buf(float* dst, __amdgpu_buffer_rsrc_t r, int pitch, int nr, int nc,
                                               int r0, int c0, int lane) {
  static_assert(GL % PS == 0, "whole rows per pass");
  constexpr int Q = PS, RPP = GL / PS;
  constexpr int MP = (Q + RPP - 1) / RPP;
  static_assert(MP * RPP == Q, "whole passes");
  const int col = lane % PS, row = lane / PS;
  float v[MP];
  if (r0 >= 0 && r0 + Q <= nr && c0 >= 0 && c0 + Q <= nc) {
    // inside the image: one offset, rows through the uniform SGPR offset
    const uint32_t o = __umul24((uint32_t)(r0 + row), (uint32_t)pitch) + (uint32_t)(c0 + col) * 4u;
#pragma unroll
    for (int m = 0; m < MP; m++) v[m] = ld_f32(r, o, m * RPP * pitch);
  } else {
    const uint32_t cb = (uint32_t)clampi(c0 + col, 0, nc - 1) * 4u;
#pragma unroll
    for (int m = 0; m < MP; m++)
      v[m] = ld_f32(r, __umul24((uint32_t)clampi(r0 + row + m * RPP, 0, nr - 1), (uint32_t)pitch) + cb, 0);
  }
#pragma unroll
  for (int m = 0; m < MP; m++) dst[(row + m * RPP) * PS + col] = v[m];
}

// Orders the wave's LDS writes before its later LDS reads (one wave, in-order
// LDS queue: only the compiler has to be kept from moving accesses).
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-lane geometry, fixed for a launch: the window pixels of the group lane
// and its gradient-patch elements (row << 8 | col).
template <int NPL>
struct lane_win {
  static constexpr int NS = lk_cfg<NPL>::NS;
  int r[NS], c[NS];
  int idx[NS];  // raster index (r + hws) * ws + (c + hws) of the window pixel
  float fr[NS], fc[NS];
  bool in[NS];
  int g[lk_lds<NPL, false>::GPASS];   // whole gradient patch
  int gf[lk_lds<NPL, false>::FPASS];  // its (ws+1)^2 core
};

template <int NPL>
__device__ __forceinline__ lane_win<NPL> make_lane_win(int ws, int lane) {
  constexpr int GL = lk_cfg<NPL>::GL, NS = lk_cfg<NPL>::NS;
  lane_win<NPL> W;
  const int hws = ws / 2, QG = ws + 3;
  const float iws = 1.f / (float)ws, iqg = 1.f / (float)QG;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    int row, col;
    if constexpr (GL == 16) {
      // a lane holds one window column at every other row: lane 8 h + c has
      // the pixels (h + 2k, c), windows up to 7x7 (8 rows, 8 columns)
      row = (lane >> 3) + 2 * k;
      col = lane & 7;
      W.in[k] = row < ws && col < ws;
    } else {
      const int i = lane + GL * k;
      row = divq(i, iws);
      col = i - row * ws;
      W.in[k] = i < ws * ws;
    }
    W.r[k] = row - hws;
    W.c[k] = col - hws;
    W.idx[k] = row * ws + col;
    W.fr[k] = (float)W.r[k];
    W.fc[k] = (float)W.c[k];
  }
#pragma unroll
  for (int m = 0; m < lk_lds<NPL, false>::GPASS; m++) {
    const int k = lane + GL * m;
    const int i = divq(k, iqg);
    W.g[m] = k < QG * QG ? (i << 8) | (k - i * QG) : -1;
  }
  const int QF = ws + 1;
  const float iqf = 1.f / (float)QF;
#pragma unroll
  for (int m = 0; m < lk_lds<NPL, false>::FPASS; m++) {
    const int k = lane + GL * m;
    const int i = divq(k, iqf);
    W.gf[m] = k < QF * QF ? ((i + 1) << 8) | (k - i * QF + 1) : -1;
  }
  return W;
}

// iterative_lk_match_5x5::match for one group of lanes.
template <typename TI, int NPL, bool SEL>
__device__ __forceinline__ match_out wave_match(const lk_level_view& L, const lk_params_dev& P,
                                                const lane_win<NPL>& W, float2 p, float2 pred_tr,
                                                lk_lds<NPL, SEL>& S, bool want_dist) {
  using LD = lk_lds<NPL, SEL>;
  constexpr int PW = LD::PW, PWA = LD::PWA;
  const match_out fail{make_float2(-1.f, -1.f), 1000.f};
  constexpr int GL = lk_cfg<NPL>::GL, NS = lk_cfg<NPL>::NS;
  const int lane = threadIdx.x & (GL - 1);
  const int ws = P.ws, hws = ws / 2, n = ws * ws;
  const int QA = ws + 5, QG = ws + 3, QB = ws + 1 + 2 * kMB;
  const float2 prediction = make_float2(add_rn(p.x, pred_tr.x), add_rn(p.y, pred_tr.y));
  float2 v = prediction;

  // ---- both patches in one batch of loads
  const int oAr = tz(p.x) - hws - 2, oAc = tz(p.y) - hws - 2;
  int oBr = tz(v.x) - hws - kMB, oBc = tz(v.y) - hws - kMB;
  wave_sync_lds();
  load_patch<TI, PWA, GL, LD::QA>(S.A, QA, L.A, L.pitchA, L.nr, L.nc, oAr, oAc, lane);
  load_patch<TI, PW, GL, LD::QB>(S.B, QB, L.B, L.pitchB, L.nr, L.nc, oBr, oBc, lane);
  wave_sync_lds();

  // ---- Sobel/8 of the previous frame at global pixel (oA + 1 + (i, j));
  // zero off the image's interior (the outer ring, and clamped reads of it).
  // A template sample (r, c) reads the gradient at its truncated coordinate
  // and one row / column further; when every sample's truncated coordinate is
  // tz(p) + (r, c) only the (ws+1)^2 core of the patch is read.
  bool dev = false;
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (W.in[k]) {
      const int xr = tz(add_rn(p.x, W.fr[k])), xc = tz(add_rn(p.y, W.fc[k]));
      dev |= xr != oAr + hws + 2 + W.r[k] || xc != oAc + hws + 2 + W.c[k];
    }
  auto sobel = [&](int code) {
    if (code < 0) return;
    const int i = code >> 8, j = code & 0xFF;
    const int yr = oAr + 1 + i, yc = oAc + 1 + j;
    float2 g = make_float2(0.f, 0.f);
    if (yr >= 1 && yr <= L.nr - 2 && yc >= 1 && yc <= L.nc - 2) {
      const float* a = &S.A[(i + 1) * PWA + (j + 1)];
      const float pmm = a[-PWA - 1], pm0 = a[-PWA], pmp = a[-PWA + 1];
      const float p0m = a[-1], p0p = a[1];
      const float ppm = a[PWA - 1], pp0 = a[PWA], ppp = a[PWA + 1];
      float gr = add_rn(ppm, mul_rn(2.f, pp0));
      gr = add_rn(gr, ppp);
      gr = sub_rn(gr, pmm);
      gr = sub_rn(gr, mul_rn(2.f, pm0));
      gr = sub_rn(gr, pmp);
      float gc = add_rn(pmp, mul_rn(2.f, p0p));
      gc = add_rn(gc, ppp);
      gc = sub_rn(gc, pmm);
      gc = sub_rn(gc, mul_rn(2.f, p0m));
      gc = sub_rn(gc, ppm);
      g = make_float2(mul_rn(gr, 0.125f), mul_rn(gc, 0.125f));
    }
    S.G[i * QG + j] = g;
  };
  if (group_ballot<GL>(dev)) {
#pragma unroll
    for (int m = 0; m < LD::GPASS; m++) sobel(W.g[m]);
  } else {
#pragma unroll
    for (int m = 0; m < LD::FPASS; m++) sobel(W.gf[m]);
  }
  wave_sync_lds();

  // ---- template: bilinear samples of A and of its gradient (pixels whose
  // truncated coordinate is inside the image)
  float2 gs[NS];
  float as[NS];
  bool sel[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const int i = W.idx[k];
    gs[k] = make_float2(0.f, 0.f);
    as[k] = 0.f;
    sel[k] = false;
    if (W.in[k]) {
      const bw w = weights(add_rn(p.x, W.fr[k]), add_rn(p.y, W.fc[k]));
      if (w.xr >= 0 && w.xr < L.nr && w.xc >= 0 && w.xc < L.nc) {
        sel[k] = true;
        const int ar = w.xr - oAr, ac = w.xc - oAc;  // in [1, ws+2]
        const float* a = &S.A[ar * PWA + ac];
        const float av = blend(w, a[0], a[PWA], a[1], a[PWA + 1]);
        as[k] = P.template_u8 ? (float)(uint8_t)(int)av : av;  // float -> uint8 through int
        const float2* g = &S.G[(ar - 1) * QG + (ac - 1)];
        const float2 g00 = g[0], g10 = g[QG], g01 = g[1], g11 = g[QG + 1];
        gs[k] = make_float2(blend(w, g00.x, g10.x, g01.x, g11.x), blend(w, g00.y, g10.y, g01.y, g11.y));
      }
      if (want_dist && i < LD::TPL) S.tpl[i] = as[k];
      if constexpr (SEL) S.mag[i] = sel[k] ? norml2(gs[k].x, gs[k].y) : -1.f;
    }
  }
  wave_sync_lds();
  if (SEL) {
    // rank among valid pixels by (|g| desc, raster index asc); keep the top_n
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int i = W.idx[k];
      if (sel[k]) {
        const float m = S.mag[i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
          const float mj = S.mag[j];
          rank += (mj > m) || (mj == m && j < i);
        }
        sel[k] = rank < P.top_n;
      }
    }
  }

  float g00 = 0.f, g01 = 0.f, g11 = 0.f;
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    if (sel[k]) {
      g00 = fmaf(gs[k].x, gs[k].x, g00);
      g01 = fmaf(gs[k].x, gs[k].y, g01);
      g11 = fmaf(gs[k].y, gs[k].y, g11);
    }
    cnt += __popcll(group_ballot<GL>(sel[k]));
  }
  const float G00 = group_sum<GL>(g00), G01 = group_sum<GL>(g01), G11 = group_sum<GL>(g11);
  const float nsel = (float)cnt;
  const float G10 = G01;
  {
    const float h = mul_rn(0.5f, add_rn(G00, G11));
    const float q = mul_rn(0.5f, sub_rn(G00, G11));
    const float rr = sqrtf(add_rn(mul_rn(q, q), mul_rn(G01, G10)));
    const float min_ev = fminf(fabsf(add_rn(h, rr)), fabsf(sub_rn(h, rr)));
    if (min_ev < P.min_eig) return fail;
  }
  const float det = sub_rn(mul_rn(G00, G11), mul_rn(G10, G01));
  const float invdet = 1.f / det;
  const float I00 = mul_rn(G11, invdet), I01 = mul_rn(-G01, invdet);
  const float I10 = mul_rn(-G10, invdet), I11 = mul_rn(G00, invdet);

  // current-frame sample at v + (r, c) from the patch
  auto sampleB = [&](int k) -> float {
    const bw w = weights(add_rn(v.x, W.fr[k]), add_rn(v.y, W.fc[k]));
    const float* b = &S.B[(w.xr - oBr) * PW + (w.xc - oBc)];
    return blend(w, b[0], b[PW], b[1], b[PW + 1]);
  };
  // the window at v has all its corners inside the patch
  auto ensure_covered = [&]() {
    const int vr = tz(v.x), vc = tz(v.y);
    if (vr - hws - 1 - oBr >= 0 && vr + hws + 2 - oBr <= QB - 1 && vc - hws - 1 - oBc >= 0 &&
        vc + hws + 2 - oBc <= QB - 1)
      return;
    oBr = vr - hws - kMB;
    oBc = vc - hws - kMB;
    wave_sync_lds();
    load_patch<TI, PW, GL, LD::QB>(S.B, QB, L.B, L.pitchB, L.nr, L.nc, oBr, oBc, lane);
    wave_sync_lds();
  };

  float nk0 = 1.f, nk1 = 1.f;
  const int rlo = 3, rhi = L.nr - 4, clo = 3, chi = L.nc - 4;
  // sqrtf(|delta|^2) >= eps and |pred - v| > 4 ws as the equivalent squared tests
  for (int it = 0; (it < P.max_iters) & (add_rn(mul_rn(nk0, nk0), mul_rn(nk1, nk1)) >= P.step_sq); it++) {
    ensure_covered();
    float b0 = 0.f, b1 = 0.f;
#pragma unroll
    for (int k = 0; k < NS; k++)
      if (sel[k]) {
        const float dt = sub_rn(as[k], sampleB(k));
        b0 = fmaf(gs[k].x, dt, b0);
        b1 = fmaf(gs[k].y, dt, b1);
      }
    b0 = group_sum<GL>(b0);
    b1 = group_sum<GL>(b1);
    nk0 = add_rn(mul_rn(I00, b0), mul_rn(I01, b1));
    nk1 = add_rn(mul_rn(I10, b0), mul_rn(I11, b1));
    v.x = add_rn(v.x, nk0);
    v.y = add_rn(v.y, nk1);
    const float dr = sub_rn(prediction.x, v.x), dc = sub_rn(prediction.y, v.y);
    if (add_rn(mul_rn(dr, dr), mul_rn(dc, dc)) > P.far_sq) return fail;
    const int vr = tz(v.x), vc = tz(v.y);
    if (!(vr >= rlo && vr <= rhi && vc >= clo && vc <= chi)) return fail;
  }
  match_out o;
  o.tr = make_float2(sub_rn(v.x, p.x), sub_rn(v.y, p.y));
  o.dist = 0.f;
  if (!want_dist) return o;

  // error with the reference's index (r+hws)*hws + (c+hws) and counter
  ensure_covered();
  float e = 0.f;
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (W.in[k]) {
      const int idx = (W.r[k] + hws) * hws + (W.c[k] + hws);
      e = add_rn(e, fabsf(sub_rn(S.tpl[idx], sampleB(k))));
    }
  const float err = group_sum<GL>(e);
  o.dist = err / (nsel + (float)n);
  return o;
}



// ------------------------------------------------------------ fast path
// Windows up to 7x7 on FP32 frames whose gradient image is given (a frame
// step computes it once per level and frame): the same operations as
// wave_match, organised around the common case of an "exact" window.
//
// A window of half width h around coordinate x is exact when every x + r,
// |r| <= h, is computed exactly: then x + r truncates to tz(x) + r and has the
// fraction of x, so all samples of the window share the bilinear weights of x
// itself (the same bits wave_match computes per sample).  For x - h >= 0 the
// sums with r <= 0 are exact, and those with r > 0 are exact when x + h stays
// below the next power of two above x, or when x's last mantissa bit is zero:
// x is then a multiple of the next binade's spacing, and x + h <= 2x never
// reaches the binade after it.  Samples then need no weights and no
// per-sample address: lane 8 h' + c reads column c of rows h' + 2k of the
// (ws+1)^2 grid of corners.  Other windows (an odd last bit near a power of
// two, near the image edge, negative coordinates) take the per-sample path.

__device__ __forceinline__ bool exact_axis(float x, int h) {
  const float lo = sub_rn(x, (float)h), hi = add_rn(x, (float)h);
  const float top = __uint_as_float(((__float_as_uint(x) >> 23) + 1u) << 23);
  // bitwise, not short-circuit: no exec-mask branches for a few compares
  return (lo >= 0.f) & ((hi < top) | ((__float_as_uint(x) & 1u) == 0u));
}

// A sufficient test for exact_axis around a start coordinate x0: every x in
// [max(2^E, h), 2^(E+1) - h), x0 in binade [2^E, 2^(E+1)), is exact (then
// x + h < 2^(E+1) exactly; the interval is empty for x0 <= 0 or NaN), and so
// is every x >= h with a zero last mantissa bit.  Tested per step instead of
// exact_axis; where it fails the per-sample path gives the same bits.
struct exact_iv {
  float lo, hi, h;
  __device__ __forceinline__ bool has(float x) const {
    return ((x >= lo) & (x < hi)) | (((__float_as_uint(x) & 1u) == 0u) & (x >= h));
  }
};
__device__ __forceinline__ exact_iv exact_interval(float x0, int h) {
  if (!(x0 > 0.f)) return exact_iv{1.f, 0.f, (float)h};
  const uint32_t e = __float_as_uint(x0) >> 23;
  return exact_iv{fmaxf(__uint_as_float(e << 23), (float)h), sub_rn(__uint_as_float((e + 1u) << 23), (float)h),
                  (float)h};
}

// Motion margin of the fast path's current-frame patch: 4 pixels make the
// patch 16 x 16, whole rows of its 16-float stride, and let the estimate move
// 3 pixels from the prediction before the patch is reloaded.
constexpr int kMB16 = 4;

struct lk_lds16 {
  static constexpr int WSMAX = 7;
  static constexpr int QB = WSMAX + 1 + 2 * kMB16;  // current-frame patch side (16)
  static constexpr int PW = 16;                     // its row stride
  static_assert(QB == PW, "whole-row patch loads");
  static constexpr int TPL = (WSMAX - 1) * (WSMAX / 2) + WSMAX;
  float B[QB * PW];
  float tpl[TPL];
};


// iterative_lk_match_5x5::match for one 16-lane group (7x7 windows, FP32
// frames, every window pixel used, gradient image in L.G).  The grid lanes
// outside the window (column 7, row 7) compute samples too, from addresses
// inside the patch or image, and their results are discarded: no lane masks
// on the exact paths.
__device__ __forceinline__ match_out wave_match16(const lk_level_view& L, const lk_params_dev& P, float2 p,
                                                  float2 pred_tr, lk_lds16& S,
                                                  bool want_dist) {
  constexpr int GL = 16, NS = lk_cfg<1>::NS, PW = lk_lds16::PW;
  constexpr int ws = 7, hws = ws / 2, n = ws * ws, QB = lk_lds16::QB;
  const match_out fail{make_float2(-1.f, -1.f), 1000.f};
  const int lane = threadIdx.x & (GL - 1);
  const int hr = lane >> 3, hc = lane & 7;  // this lane's grid row parity and column
  // the lane's window geometry, recomputed where used rather than held in
  // registers: sample k is window pixel (hr + 2k - hws, hc - hws)
  auto in_w = [&](int k) { return (hc < ws) & (hr + 2 * k < ws); };
  auto fr_w = [&](int k) { return (float)(hr + 2 * k - hws); };
  const float fc_w = (float)(hc - hws);
  const float2 prediction = make_float2(add_rn(p.x, pred_tr.x), add_rn(p.y, pred_tr.y));
  float2 v = prediction;
  const __amdgpu_buffer_rsrc_t rA = raw_rsrc(L.A), rG = raw_rsrc(L.G);

  // ---- current-frame patch around the prediction (LDS, clamped reads)
  int oBr = tz(v.x) - hws - kMB16, oBc = tz(v.y) - hws - kMB16;
  const __amdgpu_buffer_rsrc_t rB = raw_rsrc(L.B);
  wave_sync_lds();
  load_patch_buf<PW, GL>(S.B, rB, L.pitchB, L.nr, L.nc, oBr, oBc, lane);

  // ---- template: bilinear samples of A and of its gradient
  float2 gs[NS];
  float as[NS];
  bool sel[NS];
  const int pr = tz(p.x), pc = tz(p.y);
  // one row and column of margin beyond the window's corners: the grid's row 7
  // and column 7 lanes read there
  const bool exact_t = exact_axis(p.x, hws) & exact_axis(p.y, hws) & (pr + hws + 2 <= L.nr - 1) &
                       (pc + hws + 2 <= L.nc - 1);
  if (exact_t) {
    const bw w = weights(p.x, p.y);
    const uint32_t r0 = (uint32_t)(pr - hws + hr);
    const uint32_t oa0 = __umul24(r0, (uint32_t)L.pitchA) + (uint32_t)(pc - hws + hc) * 4u;
    const uint32_t og0 = __umul24(r0, (uint32_t)L.pitchG) + (uint32_t)(pc - hws + hc) * 8u;
#pragma unroll
    for (int k = 0; k < NS; k++) {
      sel[k] = in_w(k);
      // rows r0 + 2k and r0 + 2k + 1 (the latter through the uniform pitch offset)
      const uint32_t oa = oa0 + (uint32_t)(2 * k) * (uint32_t)L.pitchA;
      const uint32_t og = og0 + (uint32_t)(2 * k) * (uint32_t)L.pitchG;
      const float a00 = ld_f32(rA, oa, 0), a01 = ld_f32(rA, oa + 4, 0), a10 = ld_f32(rA, oa, L.pitchA),
                  a11 = ld_f32(rA, oa + 4, L.pitchA);
      const float2 g00 = ld_f32x2(rG, og, 0), g01 = ld_f32x2(rG, og + 8, 0), g10 = ld_f32x2(rG, og, L.pitchG),
                   g11 = ld_f32x2(rG, og + 8, L.pitchG);
      as[k] = in_w(k) ? blend(w, a00, a10, a01, a11) : 0.f;
      gs[k] = in_w(k) ? make_float2(blend(w, g00.x, g10.x, g01.x, g11.x), blend(w, g00.y, g10.y, g01.y, g11.y))
                      : make_float2(0.f, 0.f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < NS; k++) {
      gs[k] = make_float2(0.f, 0.f);
      as[k] = 0.f;
      sel[k] = false;
      if (in_w(k)) {
        const bw w = weights(add_rn(p.x, fr_w(k)), add_rn(p.y, fc_w));
        if (w.xr >= 0 && w.xr < L.nr && w.xc >= 0 && w.xc < L.nc) {
          sel[k] = true;
          const uint32_t r0 = (uint32_t)w.xr, r1 = (uint32_t)min(w.xr + 1, L.nr - 1);
          const uint32_t c0 = (uint32_t)w.xc, c1 = (uint32_t)min(w.xc + 1, L.nc - 1);
          const uint32_t ra0 = __umul24(r0, (uint32_t)L.pitchA), ra1 = __umul24(r1, (uint32_t)L.pitchA);
          const uint32_t rg0 = __umul24(r0, (uint32_t)L.pitchG), rg1 = __umul24(r1, (uint32_t)L.pitchG);
          const float a00 = ld_f32(rA, ra0 + 4 * c0, 0), a01 = ld_f32(rA, ra0 + 4 * c1, 0),
                      a10 = ld_f32(rA, ra1 + 4 * c0, 0), a11 = ld_f32(rA, ra1 + 4 * c1, 0);
          const float2 g00 = ld_f32x2(rG, rg0 + 8 * c0, 0), g01 = ld_f32x2(rG, rg0 + 8 * c1, 0),
                       g10 = ld_f32x2(rG, rg1 + 8 * c0, 0), g11 = ld_f32x2(rG, rg1 + 8 * c1, 0);
          as[k] = blend(w, a00, a10, a01, a11);
          gs[k] = make_float2(blend(w, g00.x, g10.x, g01.x, g11.x), blend(w, g00.y, g10.y, g01.y, g11.y));
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (want_dist && in_w(k) && (hr + 2 * k) * ws + hc < lk_lds16::TPL) S.tpl[(hr + 2 * k) * ws + hc] = as[k];

  // window sums with no lane masks: an unselected pixel has a zero gradient,
  // and adding fma(0, finite, acc) leaves acc (never -0 here) unchanged
  float g00 = 0.f, g01 = 0.f, g11 = 0.f;
  int cnt = n;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    g00 = fmaf(gs[k].x, gs[k].x, g00);
    g01 = fmaf(gs[k].x, gs[k].y, g01);
    g11 = fmaf(gs[k].y, gs[k].y, g11);
  }
  if (!exact_t) {
    cnt = 0;
#pragma unroll
    for (int k = 0; k < NS; k++) cnt += __popcll(group_ballot<GL>(sel[k]));
  }
  const float G00 = group_sum<GL>(g00), G01 = group_sum<GL>(g01), G11 = group_sum<GL>(g11);
  const float nsel = (float)cnt;
  const float G10 = G01;
  {
    const float h = mul_rn(0.5f, add_rn(G00, G11));
    const float q = mul_rn(0.5f, sub_rn(G00, G11));
    const float rr = sqrtf(add_rn(mul_rn(q, q), mul_rn(G01, G10)));
    const float min_ev = fminf(fabsf(add_rn(h, rr)), fabsf(sub_rn(h, rr)));
    if (min_ev < P.min_eig) return fail;
  }
  const float det = sub_rn(mul_rn(G00, G11), mul_rn(G10, G01));
  const float invdet = 1.f / det;
  const float I00 = mul_rn(G11, invdet), I01 = mul_rn(-G01, invdet);
  const float I10 = mul_rn(-G10, invdet), I11 = mul_rn(G00, invdet);

  // the window at v has all its corners inside the patch: tz(v) within
  // [cov_r, cov_r + span] (and the same for columns)
  constexpr int span = QB - 2 * hws - 4;
  int cov_r = oBr + hws + 1, cov_c = oBc + hws + 1;
  auto ensure_covered = [&]() {
    const int vr = tz(v.x), vc = tz(v.y);
    if (((uint32_t)(vr - cov_r) <= (uint32_t)span) & ((uint32_t)(vc - cov_c) <= (uint32_t)span)) return;
    oBr = vr - hws - kMB16;
    oBc = vc - hws - kMB16;
    cov_r = oBr + hws + 1;
    cov_c = oBc + hws + 1;
    wave_sync_lds();
    load_patch_buf<PW, GL>(S.B, rB, L.pitchB, L.nr, L.nc, oBr, oBc, lane);
    wave_sync_lds();
  };
  const exact_iv ivr = exact_interval(v.x, hws), ivc = exact_interval(v.y, hws);
  // current-frame samples at v + (r, c) for the lane's pixels
  auto sample_window = [&](float (&out)[NS]) {
    if (ivr.has(v.x) & ivc.has(v.y)) {
      const bw w = weights(v.x, v.y);
      const float* b0 = &S.B[(tz(v.x) - hws + hr - oBr) * PW + (tz(v.y) - hws + hc - oBc)];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const float* b = b0 + 2 * k * PW;
        out[k] = blend(w, b[0], b[PW], b[1], b[PW + 1]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < NS; k++) {
        out[k] = 0.f;
        if (in_w(k)) {
          const bw w = weights(add_rn(v.x, fr_w(k)), add_rn(v.y, fc_w));
          const float* b = &S.B[(w.xr - oBr) * PW + (w.xc - oBc)];
          out[k] = blend(w, b[0], b[PW], b[1], b[PW + 1]);
        }
      }
    }
  };

  wave_sync_lds();  // the first patch and the template prefix are in LDS
  float nk0 = 1.f, nk1 = 1.f;
  const int rlo = 3, rhi = L.nr - 4, clo = 3, chi = L.nc - 4;
  // sqrtf(|delta|^2) >= eps and |pred - v| > 4 ws as the equivalent squared tests
  for (int it = 0; (it < P.max_iters) & (add_rn(mul_rn(nk0, nk0), mul_rn(nk1, nk1)) >= P.step_sq); it++) {
    ensure_covered();
    float smp[NS];
    sample_window(smp);
    // no lane masks (zero gradients where unselected, finite samples)
    float b0 = 0.f, b1 = 0.f;
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const float dt = sub_rn(as[k], smp[k]);
      b0 = fmaf(gs[k].x, dt, b0);
      b1 = fmaf(gs[k].y, dt, b1);
    }
    b0 = group_sum<GL>(b0);
    b1 = group_sum<GL>(b1);
    nk0 = add_rn(mul_rn(I00, b0), mul_rn(I01, b1));
    nk1 = add_rn(mul_rn(I10, b0), mul_rn(I11, b1));
    v.x = add_rn(v.x, nk0);
    v.y = add_rn(v.y, nk1);
    const float dr = sub_rn(prediction.x, v.x), dc = sub_rn(prediction.y, v.y);
    // too far from the prediction, or outside domain - border: one exit
    const int vr = tz(v.x), vc = tz(v.y);
    if ((add_rn(mul_rn(dr, dr), mul_rn(dc, dc)) > P.far_sq) | ((uint32_t)(vr - rlo) > (uint32_t)(rhi - rlo)) |
        ((uint32_t)(vc - clo) > (uint32_t)(chi - clo)))
      return fail;
  }
  match_out o;
  o.tr = make_float2(sub_rn(v.x, p.x), sub_rn(v.y, p.y));
  o.dist = 0.f;
  if (!want_dist) return o;

  // error with the reference's index (r+hws)*hws + (c+hws) and counter
  ensure_covered();
  float smp[NS];
  sample_window(smp);
  float e = 0.f;
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (in_w(k)) {
      const int idx = (hr + 2 * k) * hws + hc;
      e = add_rn(e, fabsf(sub_rn(S.tpl[idx], smp[k])));
    }
  const float err = group_sum<GL>(e);
  o.dist = err / (nsel + (float)n);
  return o;
}

// Level view j read with scalar loads: the table is written before the launch
// and never during it, so it is read as constant memory (SGPRs, no LDS round
// trip and no per-field readfirstlane).
__device__ __forceinline__ lk_level_view level_view_scalar(const lk_level_view* lv, int j) {
  constexpr int NW = (int)(sizeof(lk_level_view) / 4);
  static_assert(sizeof(lk_level_view) % 4 == 0, "whole words");
  const __attribute__((address_space(4))) uint32_t* src =
      (const __attribute__((address_space(4))) uint32_t*)(lv + j);
  uint32_t w[NW];
#pragma unroll
  for (int k = 0; k < NW; k++) w[k] = src[k];
  lk_level_view out;
  __builtin_memcpy(&out, w, sizeof out);
  return out;
}

// One particle per group of lanes.  The particle at scale s and its synthetic age-1
// ancestors at p/2, p/4, ... are solved in one loop from the coarsest scale
// down (recursive_prediction / get_flow_at): an ancestor whose prediction
// leaves domain - border(3) contributes the identity, so does a match ending
// outside its level's domain, and a failed match (-1,-1) is kept; without a
// coarser scale the prediction is the speed (zero at age 1).
// FASTG: the level views carry gradient images and the window class is 7x7 on
// FP32 frames with every pixel used (wave_match16).
template <typename TI, int NPL, bool SEL, bool FASTG = false>
// Workgroup `bid` matches particles [bid * W * ppw, (bid + 1) * W * ppw), W =
// kLkWaves; the
// grid covers the capacity, so the hardware's workgroup dispatch balances the
// particles' very different costs and workgroups past the count exit at once.
__device__ inline void lk_body(const lk_args& a, int bid, int nbid, int ppw) {
  static_assert(!FASTG || (NPL == 1 && !SEL && std::is_same<TI, float>::value), "fast path: 7x7 FP32, all pixels");
  // FASTG launches are for ws == 7 (the host picks the variant)
  constexpr int GL = lk_cfg<NPL>::GL, GPW = lk_cfg<NPL>::GPW;
  using LDS = std::conditional_t<FASTG, lk_lds16, lk_lds<NPL, SEL>>;
  __shared__ LDS lds_all[kLkWaves * GPW];
  __shared__ uint32_t blk_live, blk_acc;
  const int wave = threadIdx.x >> 6;
  const int grp = (threadIdx.x & 63) / GL;
  const int lane = threadIdx.x & (GL - 1);
  LDS& S = lds_all[wave * GPW + grp];
  // Every argument is read once, before this kernel stores anything, so the
  // reads are scalar and the values stay in registers whether `a` is a kernel
  // argument or an entry of a job table; the level views are read per level
  // with scalar loads from their table.
  const lk_params_dev P = uni(a.P);
  const int s = uni(a.s), nlev = uni(a.nlevels);
  const int pd_r1 = uni(a.pd_r1), pd_r2 = uni(a.pd_r2), pd_c1 = uni(a.pd_c1), pd_c2 = uni(a.pd_c2);
  const lk_level_view* const lvp = a.lv;
  cuimg_particle_f* const parts = uni(a.parts);
  level_counters* const cnt = uni(a.cnt);
  uint32_t* const sparse = uni(a.sparse);
  const int sparse_pitch = uni(a.sparse_pitch);
  int32_t* const dec_accept = uni(a.dec_accept);
  float2* const dec_tr = uni(a.dec_tr);
  float* const dec_dist = uni(a.dec_dist);
  const uint32_t n = uni(cnt->n);
  const uint32_t per_wg = (uint32_t)kLkWaves * (uint32_t)ppw;
  const uint32_t nchunks = (n + per_wg - 1) / per_wg;
  if (bid == 0 && threadIdx.x == 0) cnt->n_matched = n;
  if ((uint32_t)bid >= nchunks) return;
  if (threadIdx.x == 0) {
    blk_live = 0;
    blk_acc = 0;
  }
  const lane_win<NPL> W = make_lane_win<NPL>(P.ws, lane);
  __syncthreads();
  // chunks bid, bid + nbid, ... of per_wg consecutive particles; the
  // workgroup's kLkWaves * GPW groups take consecutive particles of a chunk
  for (uint32_t chunk = (uint32_t)bid; chunk < nchunks; chunk += (uint32_t)nbid) {
  const uint32_t first = chunk * per_wg, last = min(n, first + per_wg);
  for (uint32_t i = first + (uint32_t)(wave * GPW + grp); i < last; i += kLkWaves * GPW) {
    cuimg_particle_f part = parts[i];
    int accept = 0;
    float2 tr = make_float2(0.f, 0.f);
    float dist = 0.f;
    const float2 pos = make_float2(part.pos_r, part.pos_c);
    if ((part.age > 0) & (tz(pos.x) >= pd_r1) & (tz(pos.x) <= pd_r2) & (tz(pos.y) >= pd_c1) & (tz(pos.y) <= pd_c2)) {
      float2 flow = make_float2(0.f, 0.f);
      for (int j = s + 1 < nlev ? nlev - 1 : s; j >= s; j--) {
        const lk_level_view lv = level_view_scalar(lvp, j);
        float2 p, ptr;
        bool run;
        if (j > s) {
          // pos / 2^(j-s): exact halvings
          p = make_float2(ldexpf(pos.x, s - j), ldexpf(pos.y, s - j));
          ptr = j == nlev - 1 ? make_float2(0.f, 0.f) : make_float2(2.f * flow.x, 2.f * flow.y);
          const int pr = tz(add_rn(p.x, ptr.x)), pc = tz(add_rn(p.y, ptr.y));
          run = (pr >= 3) & (pr <= lv.nr - 4) & (pc >= 3) & (pc <= lv.nc - 4);
        } else {
          p = pos;
          ptr = s + 1 < nlev ? make_float2(2.f * flow.x, 2.f * flow.y)
                             : (part.age == 1 ? make_float2(0.f, 0.f) : make_float2(part.speed_r, part.speed_c));
          const int pr = tz(add_rn(p.x, ptr.x)), pc = tz(add_rn(p.y, ptr.y));
          run = (pr >= pd_r1) & (pr <= pd_r2) & (pc >= pd_c1) & (pc <= pd_c2);
        }
        if (!run) {
          flow = make_float2(0.f, 0.f);
          continue;
        }
        match_out m;
        // an ancestor's match gives only its translation (the flow): its
        // error is not computed
        if constexpr (FASTG) m = wave_match16(lv, P, p, ptr, S, j == s);
        else m = wave_match<TI, NPL, SEL>(lv, P, W, p, ptr, S, j == s);
        const int mr = tz(add_rn(p.x, m.tr.x)), mc = tz(add_rn(p.y, m.tr.y));
        if (j > s) {
          flow = ((mr >= 0) & (mr < lv.nr) & (mc >= 0) & (mc < lv.nc)) ? m.tr : make_float2(0.f, 0.f);
        } else {
          const float2 pred = make_float2(add_rn(p.x, ptr.x), add_rn(p.y, ptr.y));
          tr = m.tr;
          dist = m.dist;
          if (mr >= pd_r1 && mr <= pd_r2 && mc >= pd_c1 && mc <= pd_c2 && dist < (float)P.k &&
              norml2(sub_rn((float)mr, pred.x), sub_rn((float)mc, pred.y)) < 160.f)
            accept = 1;
        }
      }
    }
    if (lane == 0) {
      if (part.age > 0) atomicAdd(&blk_live, 1u);
      if (accept) atomicAdd(&blk_acc, 1u);
      dec_accept[i] = accept;
      dec_tr[i] = tr;
      dec_dist[i] = dist;
      if (accept) {
        part.age++;
        part.speed_r = tr.x;
        part.speed_c = tr.y;
        part.pos_r = add_rn(part.pos_r, tr.x);
        part.pos_c = add_rn(part.pos_c, tr.y);
        // kernel_particle_container::move: the last writer in index order wins
        if (sparse)
          atomicMax(reinterpret_cast<int*>(px<uint32_t>(sparse, sparse_pitch, tz(part.pos_r), tz(part.pos_c))),
                    (int)i);
      } else {
        part.age = 0;
      }
      parts[i] = part;
    }
  }
  }
  __syncthreads();
  if (threadIdx.x == 0 && (blk_live | blk_acc)) {
    atomicAdd(&cnt->live_total, (unsigned long long)blk_live);
    atomicAdd(&cnt->accepted_total, (unsigned long long)blk_acc);
  }
}

template <typename T>
struct type_tag {
  using type = T;
};

// Calls f(type_tag<TI>, integral_constant<int, NPL>, bool_constant<SEL>) for the
// kernel variant of a window size, a selection size and the image type.
template <typename F>
inline void lk_variant(int ws, int top_n, int image_u8, F&& f) {
  const bool sel = top_n < ws * ws;
  auto by_npl = [&](auto ti) {
    if (ws <= lk_cfg<1>::WSMAX) {
      if (sel) f(ti, std::integral_constant<int, 1>{}, std::true_type{});
      else f(ti, std::integral_constant<int, 1>{}, std::false_type{});
    } else if (ws <= lk_cfg<2>::WSMAX) {
      if (sel) f(ti, std::integral_constant<int, 2>{}, std::true_type{});
      else f(ti, std::integral_constant<int, 2>{}, std::false_type{});
    } else {
      if (sel) f(ti, std::integral_constant<int, 4>{}, std::true_type{});
      else f(ti, std::integral_constant<int, 4>{}, std::false_type{});
    }
  };
  if (image_u8) by_npl(type_tag<uint8_t>{});
  else by_npl(type_tag<float>{});
}

}  // namespace cuimg

#pragma clang fp contract(fast)
