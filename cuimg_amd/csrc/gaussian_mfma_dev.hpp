// Separable Gaussian on the matrix cores (FP32 and FP32x4 images, wide kernels).
//
// Semantics: meta_convolve's row and column passes over taps -K..K with
// clamp-to-edge reads (cuimg/gpu/meta_convolve.h:25-128), in FP32; only the
// order in which each output's products are added differs from the VALU
// kernel (gaussian_dev.hpp), so results agree with it and with the
// reference's FP32 arithmetic to within FP32 rounding of the sums.
//
// Both passes are banded Toeplitz products on v_mfma_f32_16x16x4f32
// (D[16x16] += A[16x4] B[4x16]; lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[4(l/16)+i][l%16] in register i):
//  * row pass, per 16 x 16 block H of row-filtered values: A = 16 input rows x
//    4 input columns of the block's (16 + 2K)-column window, B = the taps, NH =
//    ceil((16 + 2K) / 4) steps (which window column a lane group takes at a
//    step is the staging layout's choice, gmfma).  A's row m is input row
//    4 (m % 4) + m / 4 of the block, which puts rows 4i .. 4i + 3 of H in
//    register i of the four lane groups;
//  * column pass, per 16 x 16 output block: B = one register of H (four
//    consecutive H rows in the four lane groups, column l % 16), A = the taps
//    for those rows, over the H blocks above, at and below the output block;
//    steps whose rows are all outside the taps are skipped at compile time
//    (10 for K = 12, 8 for K = 6).  FP32 computes the transposed product
//    (A = that H register, B = the taps), which leaves each lane with four
//    consecutive output columns of one row.
// So the row pass's D registers are the column pass's operands as they
// stand: no data movement between the passes.  The matrix cores run FP32 at
// twice the non-packed VALU rate; the wide kernels (13 and 25 taps) are
// VALU-bound on the vector path.
//
// A wave owns a strip of 64 columns (FP32: 4 column blocks) or 16 pixels
// (FP32x4: the 4 channels as 4 blocks, so a lane holds all channels of its
// output pixels and stores 16 bytes; FP32: a lane stores 16 bytes of one
// output row per column block), and a tile of rows aligned to 16.  It
// walks the tile's 16-row blocks top to bottom: the clamped input of blocks
// b + 1 and b + 2 is in flight into registers while block b, staged in LDS,
// is filtered, and output block b - 1 is written once H blocks b - 2 .. b
// exist.
#pragma once

#include "gaussian_dev.hpp"

namespace cuimg {

typedef float v4f_mfma __attribute__((ext_vector_type(4)));
typedef uint32_t v4u_t __attribute__((ext_vector_type(4)));

// Raw buffer descriptor word 3 on gfx9 (32-bit data format, no swizzle) and
// the store cache policy with the non-temporal bit (gfx940 and later).
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr int kCachePolicyNT = 2;

// Layout of the per-wave staging block for pixel type E and half width K.
// FP32: staged float x of a row is column c0 - KA + x (KA = K rounded up to 4,
// so interior strips load 16-byte chunks); the row stride RS is 2 mod 32, and
// lane group k takes window column 4 s + k at step s, so the 32 lanes of a
// half-wave read 32 different banks.  FP32x4: staged pixel x is column c0 - K
// + x, lane group k takes pixels NH k .. NH k + NH - 1.
template <typename E, int K>
struct gmfma {
  static constexpr bool RGBA = sizeof(E) == 16;
  static constexpr int NH = (16 + 2 * K + 3) / 4;  // row-pass steps
  static constexpr int SW = RGBA ? 16 : 64;        // strip width in pixels
  // FP32
  static constexpr int KA = (K + 3) / 4 * 4, OFF = KA - K;
  static constexpr int CW = (OFF + 48 + 4 * NH + 3) / 4;  // 16-byte chunks per staged row
  static constexpr int LWF = 4 * CW;                       // staged floats per row
  static constexpr int RS = LWF + ((2 - LWF % 32) + 32) % 32;
  static constexpr int NLC = (16 * CW + 63) / 64;         // chunk loads per lane per block
  // FP32x4
  static constexpr int LWP = 4 * NH;                       // staged pixels per row
  static constexpr int NL = (16 * LWP + 63) / 64;          // pixel loads per lane per block
  static constexpr int LDS_FLOATS = RGBA ? 64 * NL * 4 : 16 * RS;
};

// INTERIOR: FP32, the strip's staged window lies inside the image and the rows
// are 16-byte aligned, so every chunk is one 16-byte load (otherwise four
// clamped element loads); both types, every output column of the strip exists,
// so the stores go through a buffer descriptor bounded at the image's last row
// and need no branch.  With loads and stores free of branches the compiler's
// wait before each staging leaves the next blocks' loads in flight under the
// MFMAs; the edge instance waits more conservatively.  NT (interior only):
// non-temporal stores (gauss_job::stream_out), fixed per instance so that the
// stores stay straight-line code.
template <typename E, int K, bool INTERIOR, bool NT>
__device__ __forceinline__ void gauss_mfma_tile(const gauss_job& J, int tx, int ty, const coef_pack& cp,
                                                float* lds) {
  using G = gmfma<E, K>;
  constexpr int NH = G::NH, P = 2 * K + 1;
  const int lane = threadIdx.x;
  const int m = lane & 15, kg = lane >> 4;
  const int c0 = tx * G::SW;
  const int r0 = ty * J.th;  // th is a multiple of 16
  const int rows = min(J.th, J.nr - r0);
  const int nob = (rows + 15) >> 4;
  const int ob0 = r0 >> 4;

  // row-pass taps: B[k][n] of step s (k = lane group, n = lane % 16) is the
  // tap of window column (FP32: 4 s + k; FP32x4: NH k + s) for output column n
  float bh[NH];
#pragma unroll
  for (int s = 0; s < NH; s++) {
    const int t = (G::RGBA ? NH * kg + s : 4 * s + kg) - m;
    bh[s] = (t >= 0 && t < P) ? cp.c[t] : 0.f;
  }
  // column-pass taps: A[m][k] for H block d (-1, 0, +1 around the output
  // block) and register i: H row 16 d + 4 i + k, output row m
  float av[3][4];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int t = 16 * (d - 1) + 4 * i + kg - m + K;
      av[d][i] = (t >= 0 && t < P) ? cp.c[t] : 0.f;
    }

  // prefetch registers of one input block (rows 16 b .. 16 b + 15, clamped)
  constexpr int NPRE = G::RGBA ? G::NL : 4 * G::NLC;
  using V = std::conditional_t<G::RGBA, v4f_t, float>;
  const int s0 = c0 - G::KA;
  auto load_block = [&](int b, V (&v)[NPRE]) {
    if constexpr (G::RGBA) {
#pragma unroll
      for (int q = 0; q < G::NL; q++) {
        const int e = min(64 * q + lane, 16 * G::LWP - 1);
        const int r = e / G::LWP, x = e - r * G::LWP;
        const int gr = clampi(16 * b + r, 0, J.nr - 1), gc = clampi(c0 - K + x, 0, J.nc - 1);
        v[q] = *reinterpret_cast<const v4f_t*>(px<E>(J.in, J.in_pitch, gr, gc));
      }
    } else {
      // 16-byte chunks of the staged rows
#pragma unroll
      for (int q = 0; q < G::NLC; q++) {
        const int e = min(64 * q + lane, 16 * G::CW - 1);
        const int r = e / G::CW, cx = e - r * G::CW;
        const int gr = clampi(16 * b + r, 0, J.nr - 1), gc = s0 + 4 * cx;
        v4f_t w;
        if constexpr (INTERIOR) {
          w = *reinterpret_cast<const v4f_t*>(px<float>(J.in, J.in_pitch, gr, gc));
        } else {
#pragma unroll
          for (int k = 0; k < 4; k++) w[k] = *px<float>(J.in, J.in_pitch, gr, clampi(gc + k, 0, J.nc - 1));
        }
        v[4 * q] = w[0];
        v[4 * q + 1] = w[1];
        v[4 * q + 2] = w[2];
        v[4 * q + 3] = w[3];
      }
    }
  };
  auto stage_block = [&](const V (&v)[NPRE]) {
    if constexpr (G::RGBA) {
#pragma unroll
      for (int q = 0; q < G::NL; q++) reinterpret_cast<v4f_t*>(lds)[64 * q + lane] = v[q];  // tail: padding
    } else {
#pragma unroll
      for (int q = 0; q < G::NLC; q++) {
        const int e = 64 * q + lane;
        if (q + 1 < G::NLC || e < 16 * G::CW) {
          const int r = e / G::CW, cx = e - r * G::CW;
          float* d = lds + r * G::RS + 4 * cx;  // 8-byte aligned (RS even)
          reinterpret_cast<v2f_t*>(d)[0] = v2f_t{v[4 * q], v[4 * q + 1]};
          reinterpret_cast<v2f_t*>(d)[1] = v2f_t{v[4 * q + 2], v[4 * q + 3]};
        }
      }
    }
  };
  auto wave_lds_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  v4f_mfma h2[4], h1[4], h0[4];  // H blocks b - 2, b - 1, b (per column block / channel)
#pragma unroll
  for (int j = 0; j < 4; j++) h2[j] = h1[j] = h0[j] = v4f_mfma{0.f, 0.f, 0.f, 0.f};
  // two blocks in flight: pre[0] holds even blocks of the walk, pre[1] odd
  // ones; a block's registers are refilled with the block two further on as
  // soon as they are staged.  The refill is unconditional (the last two steps
  // reload the tile's last block, which is in L2), so at every staging the
  // other block's loads are still outstanding and the wait is only for this
  // block's
  V pre[2][NPRE];
  const int nblk = nob + 2;
  load_block(ob0 - 1, pre[0]);
  load_block(ob0, pre[1]);  // nblk >= 3
  const int arow = 4 * (m & 3) + (m >> 2);  // A row m is input row 4 (m % 4) + m / 4
  auto step = [&](int bi, V (&cur)[NPRE]) {
    const int b = ob0 - 1 + bi;
    // one wave per workgroup: ordering the wave's own LDS accesses suffices
    // (the previous block's reads before this block's writes, then the writes
    // before the reads)
    wave_lds_sync();
    stage_block(cur);
    wave_lds_sync();
    load_block(min(b + 2, ob0 + nob), cur);

    // row pass of block b
#pragma unroll
    for (int j = 0; j < 4; j++) {
      h2[j] = h1[j];
      h1[j] = h0[j];
      h0[j] = v4f_mfma{0.f, 0.f, 0.f, 0.f};
    }
    if constexpr (G::RGBA) {
      const v4f_t* row = reinterpret_cast<const v4f_t*>(lds) + arow * G::LWP + NH * kg;
#pragma unroll
      for (int s = 0; s < NH; s++) {
        const v4f_t a = row[s];
        h0[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], bh[s], h0[0], 0, 0, 0);
        h0[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], bh[s], h0[1], 0, 0, 0);
        h0[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], bh[s], h0[2], 0, 0, 0);
        h0[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], bh[s], h0[3], 0, 0, 0);
      }
    } else {
      // the four column blocks' accumulator chains interleaved step by step
      const float* row = lds + arow * G::RS + G::OFF + kg;
#pragma unroll
      for (int s = 0; s < NH; s++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          h0[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(row[16 * j + 4 * s], bh[s], h0[j], 0, 0, 0);
    }
    // the walk's first two blocks only feed H
    if constexpr (!INTERIOR) {
      if (bi < 2) return;
    }

    // column pass of output block b - 1 from H blocks b - 2, b - 1, b, the
    // four blocks' chains interleaved step by step.  FP32x4: D = C H (lane:
    // output rows 16 ob + 4 kg + i of pixel c0 + m, one channel per chain).
    // FP32: D = (C H)^T = H^T C^T; the H register that is C H's B operand is
    // H^T's A operand as it stands, and the tap register likewise, so a lane
    // ends with output row 16 ob + m, columns 16 jj + 4 kg .. 16 jj + 4 kg + 3
    // of column block jj: 16 contiguous bytes, stored without a transpose.
    const int ob = b - 1;
    v4f_mfma o[4];
#pragma unroll
    for (int j = 0; j < 4; j++) o[j] = v4f_mfma{0.f, 0.f, 0.f, 0.f};
    if (bi >= 2) {
      static_for<3>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        static_for<4>([&](auto ic) {
          constexpr int i = decltype(ic)::value;
          constexpr int lo = 16 * (d - 1) + 4 * i;  // H rows lo .. lo + 3 relative to the output block
          if constexpr (lo + 3 + K >= 0 && lo - 15 + K <= 2 * K) {
#pragma unroll
            for (int jj = 0; jj < 4; jj++) {
              const v4f_mfma& hb = d == 0 ? h2[jj] : (d == 1 ? h1[jj] : h0[jj]);
              if constexpr (G::RGBA) o[jj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[d][i], hb[i], o[jj], 0, 0, 0);
              else o[jj] = __builtin_amdgcn_mfma_f32_16x16x4f32(hb[i], av[d][i], o[jj], 0, 0, 0);
            }
          }
        });
      });
    }
    if constexpr (INTERIOR) {
      // Every output column of the strip exists.  The four stores are issued
      // at every step, so that the compiler's waits for the loads count them
      // exactly: rows past the image's last, and the whole block in the walk's
      // first two steps, lie outside the buffer descriptor and are dropped.
      const int valid = bi >= 2 ? min(16, J.nr - 16 * ob) : 0;
      const int orow = bi >= 2 ? 16 * ob : 0;
      const int pitch = J.out_pitch;
      const __amdgpu_buffer_rsrc_t out_rsrc = __builtin_amdgcn_make_buffer_rsrc(
          px<E>(J.out, pitch, orow, c0), 0, valid * pitch, kBufferRsrcWord3);
      auto store16 = [&](const v4u_t& w, int off) {
        __builtin_amdgcn_raw_buffer_store_b128(w, out_rsrc, off, 0, NT ? kCachePolicyNT : 0);
      };
      if constexpr (G::RGBA) {
        // lane: output rows 16 ob + 4 kg + i of pixel c0 + m, all four channels
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const v4f_t w = {o[0][i], o[1][i], o[2][i], o[3][i]};
          store16(__builtin_bit_cast(v4u_t, w), (4 * kg + i) * pitch + 16 * m);
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) store16(__builtin_bit_cast(v4u_t, o[jj]), m * pitch + 4 * (16 * jj + 4 * kg));
      }
      return;
    }
    if constexpr (G::RGBA) {
      // lane: output rows 16 ob + 4 kg + i of pixel c0 + m, all four channels
      const v4f_mfma o0 = o[0], o1 = o[1], o2 = o[2], o3 = o[3];
      const int c = c0 + m;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 16 * ob + 4 * kg + i;
        if (r < r0 + rows && c < J.nc) {
          float4* p = px<float4>(J.out, J.out_pitch, r, c);
          const v4f_t w = {o0[i], o1[i], o2[i], o3[i]};
          if (J.stream_out) store_nt(p, w);
          else *reinterpret_cast<v4f_t*>(p) = w;
        }
      }
    } else {
      const int r = 16 * ob + m;
      if (r < r0 + rows) {
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          const int c = c0 + 16 * jj + 4 * kg;
          const v4f_t w = o[jj];
          float* p = px<float>(J.out, J.out_pitch, r, c);
          if (J.vec_ok && c + 4 <= J.nc) {
            if (J.stream_out) store_nt(p, w);
            else *reinterpret_cast<v4f_t*>(p) = w;
          } else {
            for (int k = 0; k < 4 && c + k < J.nc; k++) p[k] = w[k];
          }
        }
      }
    }
  };
  // Steps 0 and 1 peeled and both steps of an iteration unconditional: every
  // path into a staging then has issued the same loads and stores since the
  // awaited block's loads, so the compiler's wait leaves them in flight.
  step(0, pre[0]);
  step(1, pre[1]);
  int bi = 2;
  for (; bi + 1 < nblk; bi += 2) {
    step(bi, pre[0]);
    step(bi + 1, pre[1]);
  }
  if (bi < nblk) step(bi, pre[0]);
  wait_vm<0>();  // the final refills land before the wave ends
}

}  // namespace cuimg
