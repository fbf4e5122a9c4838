// Pyramid Lucas-Kanade particle matcher for gfx950.
//
// Reference: pyrlk_match_particles_kernel (tracking2/pyrlk_tracker.hpp:676-769)
// with iterative_lk_match_5x5::match (:165-335), the bilinear sampler `li`
// (:130-146), the Sobel/8 gradient of the previous frame (:909-938) and the
// recursive coarse-to-fine prediction (predictions.h:60-74,
// pyrlk_tracker.hpp:1025-1063, transformations.h:77-80).
//
// Design: one wavefront per particle.  Lane l owns window pixels l, l+64, ...
// (raster order, up to NPL per lane).  Per level solve the lanes build the
// template (bilinear sample of the previous frame and of its Sobel gradient,
// the gradient computed on the fly from a 4x4 neighbourhood, so no gradient
// image is stored), the 2x2 normal matrix is a wave reduction (xor shuffles,
// bit-identical in every lane), and each Gauss-Newton step is one bilinear
// sample per lane plus a 2-value wave reduction; the 2x2 solve runs in FP32
// registers, redundantly in every lane, so all control flow is wave-uniform.
// No MFMA: nothing here is a dense contraction.  A particle at scale s first
// re-solves its synthetic ancestors at scales L-1 .. s+1 (as the reference's
// recursion does), each solve feeding the next one's prediction.
//
// Numerics: per-sample arithmetic (bilinear weights, samples, Sobel) is written
// without FMA contraction in the oracle's operation order, so template and
// image samples are bit-identical to the CPU oracle; only the window sums are
// reduced in tree order instead of raster order.
#include "tracker_state.hpp"

namespace cuimg {

__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float sub_rn(float a, float b) { return __fsub_rn(a, b); }

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
  return x;
}

__device__ __forceinline__ int tz(float v) { return (int)v; }

__device__ __forceinline__ float norml2(float a, float b) {
  return sqrtf(add_rn(mul_rn(a, a), mul_rn(b, b)));
}

template <typename TI>
__device__ __forceinline__ float ld(const void* img, int pitch, int r, int c) {
  return (float)*px<TI>(img, pitch, r, c);
}

// li: bilinear sample, base = coordinates truncated toward zero, clamp reads.
template <typename TI>
__device__ __forceinline__ float li(const void* img, int pitch, int nr, int nc, float qr, float qc) {
  const int xr = tz(qr), xc = tz(qc);
  const float a0 = sub_rn(qr, (float)xr), a1 = sub_rn(qc, (float)xc);
  const int r0 = clampi(xr, 0, nr - 1), r1 = clampi(xr + 1, 0, nr - 1);
  const int c0 = clampi(xc, 0, nc - 1), c1 = clampi(xc + 1, 0, nc - 1);
  const float v00 = ld<TI>(img, pitch, r0, c0), v10 = ld<TI>(img, pitch, r1, c0);
  const float v01 = ld<TI>(img, pitch, r0, c1), v11 = ld<TI>(img, pitch, r1, c1);
  const float b0 = sub_rn(1.f, a0), b1 = sub_rn(1.f, a1);
  float s = mul_rn(mul_rn(b0, b1), v00);
  s = add_rn(s, mul_rn(mul_rn(a0, b1), v10));
  s = add_rn(s, mul_rn(mul_rn(b0, a1), v01));
  s = add_rn(s, mul_rn(mul_rn(a0, a1), v11));
  return s;
}

// Sobel/8 gradient of A at pixel y (clamped); zero on the 1-pixel outer ring.
template <typename TI>
__device__ __forceinline__ float2 grad_at(const void* A, int pitch, int nr, int nc, int yr, int yc) {
  yr = clampi(yr, 0, nr - 1);
  yc = clampi(yc, 0, nc - 1);
  if (yr == 0 || yc == 0 || yr == nr - 1 || yc == nc - 1) return make_float2(0.f, 0.f);
  const float pmm = ld<TI>(A, pitch, yr - 1, yc - 1), pm0 = ld<TI>(A, pitch, yr - 1, yc),
              pmp = ld<TI>(A, pitch, yr - 1, yc + 1);
  const float p0m = ld<TI>(A, pitch, yr, yc - 1), p0p = ld<TI>(A, pitch, yr, yc + 1);
  const float ppm = ld<TI>(A, pitch, yr + 1, yc - 1), pp0 = ld<TI>(A, pitch, yr + 1, yc),
              ppp = ld<TI>(A, pitch, yr + 1, yc + 1);
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
  return make_float2(mul_rn(gr, 0.125f), mul_rn(gc, 0.125f));
}

template <typename TI>
__device__ __forceinline__ float2 li_grad(const void* A, int pitch, int nr, int nc, float qr, float qc) {
  const int xr = tz(qr), xc = tz(qc);
  const float a0 = sub_rn(qr, (float)xr), a1 = sub_rn(qc, (float)xc);
  const float b0 = sub_rn(1.f, a0), b1 = sub_rn(1.f, a1);
  const float w00 = mul_rn(b0, b1), w10 = mul_rn(a0, b1), w01 = mul_rn(b0, a1), w11 = mul_rn(a0, a1);
  const float2 g00 = grad_at<TI>(A, pitch, nr, nc, xr, xc), g10 = grad_at<TI>(A, pitch, nr, nc, xr + 1, xc);
  const float2 g01 = grad_at<TI>(A, pitch, nr, nc, xr, xc + 1),
               g11 = grad_at<TI>(A, pitch, nr, nc, xr + 1, xc + 1);
  float sr = mul_rn(w00, g00.x);
  sr = add_rn(sr, mul_rn(w10, g10.x));
  sr = add_rn(sr, mul_rn(w01, g01.x));
  sr = add_rn(sr, mul_rn(w11, g11.x));
  float sc = mul_rn(w00, g00.y);
  sc = add_rn(sc, mul_rn(w10, g10.y));
  sc = add_rn(sc, mul_rn(w01, g01.y));
  sc = add_rn(sc, mul_rn(w11, g11.y));
  return make_float2(sr, sc);
}

struct match_out {
  float2 tr;
  float dist;
};

// iterative_lk_match_5x5::match for one wave.  `lds` holds 2*64*NPL floats
// private to the wave.
template <typename TI, int NPL, bool SEL>
__device__ match_out wave_match(const lk_level_view& L, const lk_params_dev& P, float2 p, float2 pred_tr,
                                float* lds) {
  const match_out fail{make_float2(-1.f, -1.f), 1000.f};
  const int lane = threadIdx.x & 63;
  const int ws = P.ws, hws = ws / 2, n = ws * ws;
  const float2 prediction = make_float2(add_rn(p.x, pred_tr.x), add_rn(p.y, pred_tr.y));
  float2 v = prediction;
  float* tpl = lds;               // template values, raster index
  float* mag = lds + 64 * NPL;    // |gradient| for the selection

  float2 gs[NPL];
  float as[NPL];
  bool sel[NPL];
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    const int i = lane + 64 * k;
    const int r = i / ws - hws, c = i % ws - hws;
    gs[k] = make_float2(0.f, 0.f);
    as[k] = 0.f;
    sel[k] = false;
    if (i < n) {
      const float qr = add_rn(p.x, (float)r), qc = add_rn(p.y, (float)c);
      const int tr_ = tz(qr), tc_ = tz(qc);
      if (tr_ >= 0 && tr_ < L.nr && tc_ >= 0 && tc_ < L.nc) {
        sel[k] = true;
        gs[k] = li_grad<TI>(L.A, L.pitchA, L.nr, L.nc, qr, qc);
        const float a = li<TI>(L.A, L.pitchA, L.nr, L.nc, qr, qc);
        // float -> uint8 through int, as the x86 conversion of the reference
        as[k] = P.template_u8 ? (float)(uint8_t)(int)a : a;
      }
      tpl[i] = as[k];
      if (SEL) mag[i] = sel[k] ? norml2(gs[k].x, gs[k].y) : -1.f;
    }
  }
  if (SEL) {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // rank among valid pixels by (|g| desc, raster index asc); keep the top_n
#pragma unroll
    for (int k = 0; k < NPL; k++) {
      const int i = lane + 64 * k;
      if (i < n && sel[k]) {
        const float m = mag[i];
        int rank = 0;
        for (int j = 0; j < n; j++) {
          const float mj = mag[j];
          rank += (mj > m) || (mj == m && j < i);
        }
        sel[k] = rank < P.top_n;
      }
    }
  } else {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }

  float g00 = 0.f, g01 = 0.f, g11 = 0.f, cnt = 0.f;
#pragma unroll
  for (int k = 0; k < NPL; k++)
    if (sel[k]) {
      g00 += gs[k].x * gs[k].x;
      g01 += gs[k].x * gs[k].y;
      g11 += gs[k].y * gs[k].y;
      cnt += 1.f;
    }
  const float G00 = wave_sum(g00), G01 = wave_sum(g01), G11 = wave_sum(g11);
  const float nsel = wave_sum(cnt);
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

  float nk0 = 1.f, nk1 = 1.f;
  const int rlo = 3, rhi = L.nr - 4, clo = 3, chi = L.nc - 4;
  for (int it = 0; it < P.max_iters && sqrtf(add_rn(mul_rn(nk0, nk0), mul_rn(nk1, nk1))) >= P.eps; it++) {
    float b0 = 0.f, b1 = 0.f;
#pragma unroll
    for (int k = 0; k < NPL; k++)
      if (sel[k]) {
        const int i = lane + 64 * k;
        const int r = i / ws - hws, c = i % ws - hws;
        const float dt =
            sub_rn(as[k], li<TI>(L.B, L.pitchB, L.nr, L.nc, add_rn(v.x, (float)r), add_rn(v.y, (float)c)));
        b0 += gs[k].x * dt;
        b1 += gs[k].y * dt;
      }
    b0 = wave_sum(b0);
    b1 = wave_sum(b1);
    nk0 = add_rn(mul_rn(I00, b0), mul_rn(I01, b1));
    nk1 = add_rn(mul_rn(I10, b0), mul_rn(I11, b1));
    v.x = add_rn(v.x, nk0);
    v.y = add_rn(v.y, nk1);
    if (norml2(sub_rn(prediction.x, v.x), sub_rn(prediction.y, v.y)) > (float)(ws * 4)) return fail;
    const int vr = tz(v.x), vc = tz(v.y);
    if (!(vr >= rlo && vr <= rhi && vc >= clo && vc <= chi)) return fail;
  }

  // error with the reference's index (r+hws)*hws + (c+hws) and counter
  float e = 0.f;
#pragma unroll
  for (int k = 0; k < NPL; k++) {
    const int i = lane + 64 * k;
    if (i < n) {
      const int r = i / ws - hws, c = i % ws - hws;
      const int idx = (r + hws) * hws + (c + hws);
      e += fabsf(sub_rn(tpl[idx], li<TI>(L.B, L.pitchB, L.nr, L.nc, add_rn(v.x, (float)r), add_rn(v.y, (float)c))));
    }
  }
  const float err = wave_sum(e);
  match_out o;
  o.tr = make_float2(sub_rn(v.x, p.x), sub_rn(v.y, p.y));
  o.dist = err / (nsel + (float)n);
  __builtin_amdgcn_wave_barrier();
  return o;
}

// recursive_prediction for a particle of scale s: the synthetic age-1 ancestors
// at p/2, p/4, ... are matched from the coarsest scale down; a scale whose
// prediction leaves domain - border(3) contributes the identity, a match ending
// outside the domain contributes the identity, a failed match (-1,-1) is kept.
template <typename TI, int NPL, bool SEL>
__device__ float2 prediction(const lk_args& a, float2 pos, int age, float2 speed, float* lds) {
  const int s = a.s, L = a.nlevels;
  if (s + 1 >= L) return age == 1 ? make_float2(0.f, 0.f) : speed;
  float2 flow = make_float2(0.f, 0.f);
  for (int j = L - 1; j > s; j--) {
    const lk_level_view& lv = a.lv[j];
    const float2 ptr = j == L - 1 ? make_float2(0.f, 0.f) : make_float2(2.f * flow.x, 2.f * flow.y);
    // pos / 2^(j-s): repeated exact halvings
    const float2 p = make_float2(ldexpf(pos.x, s - j), ldexpf(pos.y, s - j));
    const float2 pred = make_float2(add_rn(p.x, ptr.x), add_rn(p.y, ptr.y));
    const int pr = tz(pred.x), pc = tz(pred.y);
    if (pr >= 3 && pr <= lv.nr - 4 && pc >= 3 && pc <= lv.nc - 4) {
      const match_out m = wave_match<TI, NPL, SEL>(lv, a.P, p, ptr, lds);
      const int mr = tz(add_rn(p.x, m.tr.x)), mc = tz(add_rn(p.y, m.tr.y));
      flow = (mr >= 0 && mr < lv.nr && mc >= 0 && mc < lv.nc) ? m.tr : make_float2(0.f, 0.f);
    } else {
      flow = make_float2(0.f, 0.f);
    }
  }
  return make_float2(2.f * flow.x, 2.f * flow.y);
}

template <typename TI, int NPL, bool SEL>
__global__ __launch_bounds__(256) void lk_kernel(lk_args a) {
  __shared__ float lds_all[4][2 * 64 * NPL];
  __shared__ uint32_t blk_live, blk_acc;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  float* lds = lds_all[wave];
  const uint32_t n = a.cnt->n;
  const int s = a.s;
  const lk_level_view& lv = a.lv[s];
  if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt->n_matched = n;
  if (threadIdx.x == 0) {
    blk_live = 0;
    blk_acc = 0;
  }
  __syncthreads();
  for (uint32_t i = blockIdx.x * 4 + wave; i < n; i += gridDim.x * 4) {
    cuimg_particle_f part = a.parts[i];
    int accept = 0;
    float2 tr = make_float2(0.f, 0.f);
    float dist = 0.f;
    const float2 pos = make_float2(part.pos_r, part.pos_c);
    const int r = tz(pos.x), c = tz(pos.y);
    if (part.age > 0 && r >= a.pd_r1 && r <= a.pd_r2 && c >= a.pd_c1 && c <= a.pd_c2) {
      const float2 ptr = prediction<TI, NPL, SEL>(a, pos, part.age, make_float2(part.speed_r, part.speed_c), lds);
      const float2 pred = make_float2(add_rn(pos.x, ptr.x), add_rn(pos.y, ptr.y));
      const int pr = tz(pred.x), pc = tz(pred.y);
      if (pr >= a.pd_r1 && pr <= a.pd_r2 && pc >= a.pd_c1 && pc <= a.pd_c2) {
        const match_out m = wave_match<TI, NPL, SEL>(lv, a.P, pos, ptr, lds);
        tr = m.tr;
        dist = m.dist;
        const int mr = tz(add_rn(pos.x, tr.x)), mc = tz(add_rn(pos.y, tr.y));
        if (mr >= a.pd_r1 && mr <= a.pd_r2 && mc >= a.pd_c1 && mc <= a.pd_c2 && dist < (float)a.P.k &&
            norml2(sub_rn((float)mr, pred.x), sub_rn((float)mc, pred.y)) < 160.f)
          accept = 1;
      }
    }
    if (lane == 0) {
      if (part.age > 0) atomicAdd(&blk_live, 1u);
      if (accept) atomicAdd(&blk_acc, 1u);
      a.dec_accept[i] = accept;
      a.dec_tr[i] = tr;
      a.dec_dist[i] = dist;
      if (accept) {
        part.age++;
        part.speed_r = tr.x;
        part.speed_c = tr.y;
        part.pos_r = add_rn(part.pos_r, tr.x);
        part.pos_c = add_rn(part.pos_c, tr.y);
        // kernel_particle_container::move: the last writer in index order wins
        atomicMax(reinterpret_cast<int*>(px<uint32_t>(a.sparse, a.sparse_pitch, tz(part.pos_r), tz(part.pos_c))),
                  (int)i);
      } else {
        part.age = 0;
      }
      a.parts[i] = part;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && (blk_live | blk_acc)) {
    atomicAdd(&a.cnt->live_total, (unsigned long long)blk_live);
    atomicAdd(&a.cnt->accepted_total, (unsigned long long)blk_acc);
  }
}

template <typename TI>
static int launch_ti(cuimg_hip_ctx* ctx, const lk_args& a, int blocks) {
  const int n = a.P.ws * a.P.ws;
  const bool sel = a.P.top_n < n;
  if (n <= 64) {
    if (sel) lk_kernel<TI, 1, true><<<blocks, 256, 0, ctx->stream>>>(a);
    else lk_kernel<TI, 1, false><<<blocks, 256, 0, ctx->stream>>>(a);
  } else if (n <= 128) {
    if (sel) lk_kernel<TI, 2, true><<<blocks, 256, 0, ctx->stream>>>(a);
    else lk_kernel<TI, 2, false><<<blocks, 256, 0, ctx->stream>>>(a);
  } else {
    if (sel) lk_kernel<TI, 4, true><<<blocks, 256, 0, ctx->stream>>>(a);
    else lk_kernel<TI, 4, false><<<blocks, 256, 0, ctx->stream>>>(a);
  }
  return 0;
}

int launch_lk(cuimg_hip_ctx* ctx, const lk_args& a, int image_u8) {
  if (a.P.ws < 1 || a.P.ws > 15 || (a.P.ws & 1) == 0)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk: window %d not supported (odd, <= 15)", a.P.ws);
  const int blocks = 2048;  // 8192 waves, grid-stride over the device-side count
  return launch(ctx, "lk_track", [&] {
    if (image_u8) launch_ti<uint8_t>(ctx, a, blocks);
    else launch_ti<float>(ctx, a, blocks);
  });
}

}  // namespace cuimg
