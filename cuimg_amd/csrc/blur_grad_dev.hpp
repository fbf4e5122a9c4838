// Blurred frame and its gradient in one pass (the blurred-frame matcher
// profile of the tracker's frame step).
//
// Per level of the new frame: the separable Gaussian of the u8 level
// (gaussian_dev.hpp, meta_convolve semantics, cuimg/gpu/meta_convolve.h:25-128)
// and the Sobel/8 gradient of that blurred level (pyrlk_cpu::match_particles,
// tracking2/pyrlk_tracker.hpp:909-938, taken of the blurred level, zero on the
// 1-pixel outer ring): g_r = (I(r+1,c-1) + 2 I(r+1,c) + I(r+1,c+1) - I(r-1,c-1)
// - 2 I(r-1,c) - I(r-1,c+1)) / 8, each operation rounded in that order, and
// g_c likewise over columns.  The blurred values are gauss_tile's bits.
//
// A wave owns a strip of 256 columns and a tile of TH output rows, as
// gauss_tile, and produces the blurred rows r0 - 1 .. r_end (two more than it
// stores) so that the Sobel of every stored row finds its rows above and
// below.  Each lane filters its 4 columns plus one extra column (lane 0:
// c0 - 1, lane 63: c0 + 256; the other lanes' extra value is unused), so the
// LDS row carries K + 1 halo columns per side; a column's neighbours inside
// the strip come from the adjacent lanes.  Blurred values at columns or rows
// outside the domain (clamped reads) only feed ring pixels, whose gradient
// is zero.
#pragma once

#include "gaussian_dev.hpp"

#pragma clang fp contract(off)

namespace cuimg {

template <int K>
__device__ __forceinline__ void gauss_sobel_tile(const gauss_job& J, float2* gout, int gpitch, int tx, int ty,
                                                 const coef_pack& cp) {
  constexpr int VEC = 4, W = 64 * VEC, P = 2 * K + 1, H = K + 1, LW = W + 2 * H;
  __shared__ __align__(16) float lds[2][LW];  // column c0 - H + x at index x
  const int lane = threadIdx.x;
  const int c0 = tx * W;
  const int r0 = ty * J.th;
  const int r_end = min(r0 + J.th, J.nr);
  const int first = r0 - 1 - K, last = r_end + K;  // input rows
  const int col = c0 + VEC * lane;
  const bool vec_in = J.vec_ok && col + VEC <= J.nc;
  // the lane's extra column, as an LDS offset of its first tap
  const int xoff = lane == 0 ? 0 : W + 1;

  float c[P];
#pragma unroll
  for (int m = 0; m < P; m++) c[m] = cp.c[m];

  auto load_row = [&](int ri, float (&v)[VEC], float& hl, float& hr) {
    const uint8_t* row = px<uint8_t>(J.in, J.in_pitch, clampi(ri, 0, J.nr - 1), 0);
    if (vec_in) {
      const uchar4 q = *reinterpret_cast<const uchar4*>(row + col);
      v[0] = q.x;
      v[1] = q.y;
      v[2] = q.z;
      v[3] = q.w;
    } else {
#pragma unroll
      for (int i = 0; i < VEC; i++) v[i] = row[min(col + i, J.nc - 1)];
    }
    if (lane < H) {
      hl = row[max(c0 - H + lane, 0)];
      hr = row[min(c0 + W + lane, J.nc - 1)];
    }
  };

  float hist[P][VEC + 1];
  float cur[VEC], hl = 0.f, hr = 0.f;
  load_row(first, cur, hl, hr);
  int buf = 0;
  // blurred rows o - 2 (up) and o - 1 (mid): [left neighbour, 4 columns, right neighbour]
  float up[VEC + 2], mid[VEC + 2];
#pragma unroll
  for (int i = 0; i < VEC + 2; i++) up[i] = mid[i] = 0.f;

  for (int base = first; base <= last; base += P) {
    static_for<P>([&](auto tc) {
      constexpr int t = decltype(tc)::value;
      const int ri = base + t;
      if (ri > last) return;
#pragma unroll
      for (int i = 0; i < VEC; i++) lds[buf][H + VEC * lane + i] = cur[i];
      if (lane < H) {
        lds[buf][lane] = hl;
        lds[buf][H + W + lane] = hr;
      }
      if (ri + 1 <= last) load_row(ri + 1, cur, hl, hr);
      __syncthreads();

      // horizontal taps: the lane's columns col + i read LDS 4 lane + 1 + i + m
      float x[VEC + 2 * K];
#pragma unroll
      for (int m = 0; m < VEC + 2 * K; m++) x[m] = lds[buf][VEC * lane + 1 + m];
#pragma unroll
      for (int i = 0; i < VEC; i++) {
        float s = mul_e(c[2 * K], x[i + 2 * K]);
#pragma unroll
        for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], x[i + m], s);
        hist[t][i] = s;
      }
      {
        float s = mul_e(c[2 * K], lds[buf][xoff + 2 * K]);
#pragma unroll
        for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], lds[buf][xoff + m], s);
        hist[t][VEC] = s;
      }
      buf ^= 1;

      const int o = ri - K;  // the blurred row the ring now completes
      if (o >= r0 - 1) {
        float dn[VEC + 2];
        float e;
#pragma unroll
        for (int i = 0; i <= VEC; i++) {
          float s = mul_e(c[2 * K], hist[t][i]);
#pragma unroll
          for (int m = 2 * K - 1; m >= 0; m--) s = fma_e(c[m], hist[(t + m + 1) % P][i], s);
          if (i < VEC) dn[1 + i] = s;
          else e = s;
        }
        const float left = __shfl_up(dn[VEC], 1, 64), right = __shfl_down(dn[1], 1, 64);
        dn[0] = lane == 0 ? e : left;
        dn[VEC + 1] = lane == 63 ? e : right;
        if (o >= r0 && o < r_end) {
          float* orow = px<float>(J.out, J.out_pitch, o, 0);
          if (J.vec_ok && col + VEC <= J.nc) {
            *reinterpret_cast<float4*>(orow + col) = make_float4(dn[1], dn[2], dn[3], dn[4]);
          } else {
#pragma unroll
            for (int i = 0; i < VEC; i++)
              if (col + i < J.nc) orow[col + i] = dn[1 + i];
          }
        }
        const int g = o - 1;  // gradient row: blurred rows g - 1 (up), g (mid), g + 1 (dn)
        if (g >= r0 && g < r_end) {
          float2 gv[VEC];
#pragma unroll
          for (int j = 0; j < VEC; j++) {
            const int cc = col + j;
            gv[j] = make_float2(0.f, 0.f);
            if (g > 0 && g < J.nr - 1 && cc > 0 && cc < J.nc - 1) {
              const float pmm = up[j], pm0 = up[j + 1], pmp = up[j + 2];
              const float p0m = mid[j], p0p = mid[j + 2];
              const float ppm = dn[j], pp0 = dn[j + 1], ppp = dn[j + 2];
              float gr = ppm + 2.f * pp0;
              gr = gr + ppp;
              gr = gr - pmm;
              gr = gr - 2.f * pm0;
              gr = gr - pmp;
              float gc = pmp + 2.f * p0p;
              gc = gc + ppp;
              gc = gc - pmm;
              gc = gc - 2.f * p0m;
              gc = gc - ppm;
              gv[j] = make_float2(gr * 0.125f, gc * 0.125f);
            }
          }
          float2* grow = px<float2>(gout, gpitch, g, 0);
          if (col + VEC <= J.nc) {
            reinterpret_cast<float4*>(grow + col)[0] = make_float4(gv[0].x, gv[0].y, gv[1].x, gv[1].y);
            reinterpret_cast<float4*>(grow + col)[1] = make_float4(gv[2].x, gv[2].y, gv[3].x, gv[3].y);
          } else {
#pragma unroll
            for (int j = 0; j < VEC; j++)
              if (col + j < J.nc) grow[col + j] = gv[j];
          }
        }
#pragma unroll
        for (int i = 0; i < VEC + 2; i++) {
          up[i] = mid[i];
          mid[i] = dn[i];
        }
      }
    });
  }
}

}  // namespace cuimg

#pragma clang fp contract(fast)
