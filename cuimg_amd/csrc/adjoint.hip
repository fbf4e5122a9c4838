// Adjoint of the blurred photometric LK loss, for gfx950.
//
//   L = sum_p sum_w ( (G*I1)(x_p + d_p + w) - (G*I0)(x_p + w) )^2
//
// G*I is the clamp-to-edge separable Gaussian of cuimg_hip_gaussian_sep (taps
// -K..K, sums from tap +K down), samples are the matcher's bilinear `li`
// (tracking2/pyrlk_tracker.hpp:130-146: base pixel truncated toward zero,
// clamped corners), w runs over a ws x ws window.  This is the photometric
// term the reference matcher drives to zero (pyrlk_tracker.hpp:283-317); the
// loss and its adjoint are an addition on top of the reference.
//
//   dL/dd_p  = sum_w 2 r ∇(bilinear of G*I1)        (per particle)
//   dL/dI1   = G^T * S,  S = sum_p sum_w 2 r · bilinear weights (splat)
//
// Design: one wave per particle.  The blurred values are never stored: each
// wave loads the (ws+1+2K)^2 patch of I1 (and of I0) around its window into
// LDS, recomputes the (ws+1)^2 blurred patch (row then column pass, the same
// tap order as the forward kernel), samples it, reduces the loss and dL/dd
// with wave shuffles, and gathers the splat of each blurred pixel from the
// <= 4 window samples touching it (all samples of a particle share one
// sub-pixel offset), adding it to S with one atomic per blurred pixel.
// A second kernel applies G^T to S with the adjoint of the clamp: taps that
// read past an edge accumulate into the edge pixel.  Particles whose window
// touches the image edge take a direct, clamped path for both.
#include "common.hpp"

namespace cuimg {

constexpr int kAdjMaxWs = 7;
constexpr int kAdjMaxK = 6;

struct adj_args {
  const float* I0;
  const float* I1;
  int pitch0, pitch1;  // bytes
  int nr, nc;
  const float2* pos;
  const float2* d;
  uint32_t n;
  int ws, K;
  float c[2 * kAdjMaxK + 1];
  float* S;            // nr*nc dense accumulator (zeroed by the host code)
  float* loss_p;       // per particle
  float2* dL_dd;
  double* loss_total;   // FP64 accumulation of the per-particle sums
};

__device__ __forceinline__ float ldI(const float* I, int pitch, int nr, int nc, int r, int c) {
  return *px<float>(I, pitch, clampi(r, 0, nr - 1), clampi(c, 0, nc - 1));
}

// Blurred value at a (clamped) pixel by the direct separable sum.
__device__ float blur_at(const adj_args& a, const float* I, int pitch, int yr, int yc) {
  yr = clampi(yr, 0, a.nr - 1);
  yc = clampi(yc, 0, a.nc - 1);
  const int K = a.K;
  float col = 0.f;
  for (int i = K; i >= -K; i--) {
    float row = 0.f;
    for (int j = K; j >= -K; j--) row = fmaf(a.c[j + K], ldI(I, pitch, a.nr, a.nc, yr + i, yc + j), row);
    col = fmaf(a.c[i + K], row, col);
  }
  return col;
}

__device__ __forceinline__ float wsum(float x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
  return x;
}

// Blurred (ws+1)^2 patch whose pixel (0,0) is global pixel (br, bc), computed
// from the I patch staged in LDS.  Lane l computes patch pixel l (row-major).
__device__ float blurred_patch_value(const adj_args& a, const float* Ipatch, float* rowtmp, int P, int Q, int lane) {
  // Ipatch: Q x Q (Q = P + 2K) raw pixels; rowtmp: Q x P row-filtered
  const int K = a.K;
  for (int k = lane; k < Q * P; k += 64) {
    const int r = k / P, c = k % P;
    float s = 0.f;
    for (int j = K; j >= -K; j--) s = fmaf(a.c[j + K], Ipatch[r * Q + c + K + j], s);
    rowtmp[k] = s;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  float v = 0.f;
  if (lane < P * P) {
    const int r = lane / P, c = lane % P;
    for (int i = K; i >= -K; i--) v = fmaf(a.c[i + K], rowtmp[(r + K + i) * P + c], v);
  }
  return v;
}

__global__ __launch_bounds__(256) void lk_adjoint_kernel(adj_args a) {
  constexpr int PMAX = kAdjMaxWs + 1;
  constexpr int QMAX = PMAX + 2 * kAdjMaxK;
  __shared__ float s_ip[4][2][QMAX * QMAX];
  __shared__ double s_loss[4];
  __shared__ float s_rt[4][QMAX * PMAX];
  __shared__ float s_b1[4][PMAX * PMAX];
  __shared__ float s_b0[4][PMAX * PMAX];
  __shared__ float s_t[4][PMAX * PMAX];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ws = a.ws, hws = ws / 2, K = a.K, P = ws + 1, Q = P + 2 * K;
  double wave_loss = 0.0;  // lane 0: this wave's particles, added once per workgroup at the end
  for (uint32_t p = blockIdx.x * 4 + wave; p < a.n; p += gridDim.x * 4) {
    const float2 x = a.pos[p], d = a.d[p];
    const float q1r = x.x + d.x, q1c = x.y + d.y;
    // window sample (r,c) sits at q + (r,c); every sample shares the fraction of q
    const int t1r = (int)q1r, t1c = (int)q1c, t0r = (int)x.x, t0c = (int)x.y;
    const float a1r = q1r - (float)t1r, a1c = q1c - (float)t1c;
    const float a0r = x.x - (float)t0r, a0c = x.y - (float)t0c;
    const int b1r = t1r - hws, b1c = t1c - hws, b0r = t0r - hws, b0c = t0c - hws;
    const bool interior = q1r >= 0.f && q1c >= 0.f && x.x >= 0.f && x.y >= 0.f && b1r >= 0 && b1c >= 0 &&
                          b0r >= 0 && b0c >= 0 && b1r + P <= a.nr && b1c + P <= a.nc && b0r + P <= a.nr &&
                          b0c + P <= a.nc;
    float* B1 = s_b1[wave];
    float* B0 = s_b0[wave];
    if (!interior) {
      // window touching an edge: every sample on its own (truncation toward
      // zero of each coordinate, clamped corners, blurred values summed directly)
      float lp = 0.f, gr = 0.f, gc = 0.f;
      const int n = ws * ws;
      if (lane < n) {
        const int r = lane / ws - hws, c = lane % ws - hws;
        const float qr = q1r + (float)r, qc = q1c + (float)c, pr = x.x + (float)r, pc = x.y + (float)c;
        const int xr = (int)qr, xc = (int)qc, yr = (int)pr, yc = (int)pc;
        const float ar = qr - (float)xr, ac = qc - (float)xc, er = pr - (float)yr, ec = pc - (float)yc;
        const int r0 = clampi(xr, 0, a.nr - 1), r1 = clampi(xr + 1, 0, a.nr - 1);
        const int c0 = clampi(xc, 0, a.nc - 1), c1 = clampi(xc + 1, 0, a.nc - 1);
        const float v00 = blur_at(a, a.I1, a.pitch1, r0, c0), v10 = blur_at(a, a.I1, a.pitch1, r1, c0);
        const float v01 = blur_at(a, a.I1, a.pitch1, r0, c1), v11 = blur_at(a, a.I1, a.pitch1, r1, c1);
        const float u00 = blur_at(a, a.I0, a.pitch0, yr, yc), u10 = blur_at(a, a.I0, a.pitch0, yr + 1, yc);
        const float u01 = blur_at(a, a.I0, a.pitch0, yr, yc + 1), u11 = blur_at(a, a.I0, a.pitch0, yr + 1, yc + 1);
        const float v1 = (1 - ar) * (1 - ac) * v00 + ar * (1 - ac) * v10 + (1 - ar) * ac * v01 + ar * ac * v11;
        const float v0 = (1 - er) * (1 - ec) * u00 + er * (1 - ec) * u10 + (1 - er) * ec * u01 + er * ec * u11;
        const float res = v1 - v0;
        const float t = 2.f * res;
        lp = res * res;
        gr = t * ((1 - ac) * (v10 - v00) + ac * (v11 - v01));
        gc = t * ((1 - ar) * (v01 - v00) + ar * (v11 - v10));
        atomicAdd(&a.S[(size_t)r0 * a.nc + c0], t * (1 - ar) * (1 - ac));
        atomicAdd(&a.S[(size_t)r1 * a.nc + c0], t * ar * (1 - ac));
        atomicAdd(&a.S[(size_t)r0 * a.nc + c1], t * (1 - ar) * ac);
        atomicAdd(&a.S[(size_t)r1 * a.nc + c1], t * ar * ac);
      }
      lp = wsum(lp);
      gr = wsum(gr);
      gc = wsum(gc);
      if (lane == 0) {
        a.loss_p[p] = lp;
        a.dL_dd[p] = make_float2(gr, gc);
        wave_loss += (double)lp;
      }
      continue;
    }
    {
      // both raw patches in one batch of loads, all issued before the first LDS write
      constexpr int NQ = (QMAX * QMAX + 63) / 64;
      float raw[2][NQ];
#pragma unroll
      for (int img = 0; img < 2; img++) {
        const float* I = img ? a.I1 : a.I0;
        const int pitch = img ? a.pitch1 : a.pitch0;
        const int br = img ? b1r : b0r, bc = img ? b1c : b0c;
#pragma unroll
        for (int m = 0; m < NQ; m++) {
          const int k = lane + 64 * m;
          if (k < Q * Q) raw[img][m] = ldI(I, pitch, a.nr, a.nc, br - K + k / Q, bc - K + k % Q);
        }
      }
#pragma unroll
      for (int img = 0; img < 2; img++)
#pragma unroll
        for (int m = 0; m < NQ; m++) {
          const int k = lane + 64 * m;
          if (k < Q * Q) s_ip[wave][img][k] = raw[img][m];
        }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      for (int img = 0; img < 2; img++) {
        const float v = blurred_patch_value(a, s_ip[wave][img], s_rt[wave], P, Q, lane);
        if (lane < P * P) (img ? B1 : B0)[lane] = v;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    }

    float lp = 0.f, gr = 0.f, gc = 0.f, t = 0.f;
    const int n = ws * ws;
    if (lane < n) {
      const int r = lane / ws, c = lane % ws;  // window offset (r - hws, c - hws)
      const float v00 = B1[r * P + c], v10 = B1[(r + 1) * P + c], v01 = B1[r * P + c + 1],
                  v11 = B1[(r + 1) * P + c + 1];
      const float u00 = B0[r * P + c], u10 = B0[(r + 1) * P + c], u01 = B0[r * P + c + 1],
                  u11 = B0[(r + 1) * P + c + 1];
      const float v1 = (1 - a1r) * (1 - a1c) * v00 + a1r * (1 - a1c) * v10 + (1 - a1r) * a1c * v01 + a1r * a1c * v11;
      const float v0 = (1 - a0r) * (1 - a0c) * u00 + a0r * (1 - a0c) * u10 + (1 - a0r) * a0c * u01 + a0r * a0c * u11;
      const float res = v1 - v0;
      lp = res * res;
      t = 2.f * res;
      gr = t * ((1 - a1c) * (v10 - v00) + a1c * (v11 - v01));
      gc = t * ((1 - a1r) * (v01 - v00) + a1r * (v11 - v10));
    }
    lp = wsum(lp);
    gr = wsum(gr);
    gc = wsum(gc);
    if (lane == 0) {
      a.loss_p[p] = lp;
      a.dL_dd[p] = make_float2(gr, gc);
      wave_loss += (double)lp;
    }
    // splat 2r with the bilinear weights onto the blurred pixels
    const float w00 = (1 - a1r) * (1 - a1c), w10 = a1r * (1 - a1c), w01 = (1 - a1r) * a1c, w11 = a1r * a1c;
    {
      float* T = s_t[wave];
      if (lane < n) T[lane] = t;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (lane < P * P) {
        const int i = lane / P, j = lane % P;
        float s = 0.f;
        if (i < ws && j < ws) s += w00 * T[i * ws + j];
        if (i >= 1 && j < ws) s += w10 * T[(i - 1) * ws + j];
        if (i < ws && j >= 1) s += w01 * T[i * ws + j - 1];
        if (i >= 1 && j >= 1) s += w11 * T[(i - 1) * ws + j - 1];
        if (s != 0.f) atomicAdd(&a.S[(size_t)(b1r + i) * a.nc + (b1c + j)], s);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // one FP64 atomic per workgroup for the total loss
  if (lane == 0) s_loss[wave] = wave_loss;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (s_loss[0] + s_loss[1]) + (s_loss[2] + s_loss[3]);
    if (t != 0.0) atomicAdd(a.loss_total, t);
  }
}

// One pass of G^T along columns (axis 1) or rows (axis 0) with the clamp
// adjoint; in and out are dense nr x nc (out with an explicit pitch).
__global__ __launch_bounds__(256) void blur_adjoint_pass(const float* in, int in_pitch, float* out, int out_pitch,
                                                         int nr, int nc, int axis, int K, adj_args a) {
  const int total = nr * nc;  // < 2^31 (checked by the host)
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int r = idx / nc, c = idx - r * nc;
    const int y = axis ? c : r, N = axis ? nc : nr;
    auto at = [&](int x) { return axis ? *px<float>(in, in_pitch, r, x) : *px<float>(in, in_pitch, x, c); };
    float s = 0.f;
    for (int x = max(0, y - K); x <= min(N - 1, y + K); x++) s = fmaf(a.c[y - x + K], at(x), s);
    if (y == 0)
      for (int x = 0; x < min(K, N); x++)
        for (int j = -K; j <= -x - 1; j++) s = fmaf(a.c[j + K], at(x), s);
    if (y == N - 1)
      for (int x = max(0, N - K); x < N; x++)
        for (int j = N - x; j <= K; j++) s = fmaf(a.c[j + K], at(x), s);
    *px<float>(out, out_pitch, r, c) = s;
  }
}

__global__ void zero_kernel(float* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

int lk_loss_adjoint(cuimg_hip_ctx* ctx, const cuimg_img* I0, const cuimg_img* I1, const float* pos2,
                    const float* d2, uint32_t n, int sigma, int K, int ws, double* loss_total, float* loss_p,
                    float* dL_dd2, const cuimg_img* dL_dI1, float* scratch, size_t scratch_floats) {
  if (!ctx) return CUIMG_E_INVALID;
  if (!loss_total || !loss_p || !dL_dd2 || !scratch || (n && (!pos2 || !d2)))
    return set_error(ctx, CUIMG_E_INVALID, "lk_loss_adjoint: null output, scratch or particle array");
  if (!I0 || !I1 || !I0->begin || !I1->begin || (dL_dI1 && !dL_dI1->begin))
    return set_error(ctx, CUIMG_E_INVALID, "lk_loss_adjoint: null image");
  if (I0->dtype != CUIMG_F32 || I1->dtype != CUIMG_F32 || I0->nrows != I1->nrows ||
      I0->ncols != I1->ncols)
    return set_error(ctx, CUIMG_E_INVALID, "lk_loss_adjoint: I0, I1 must be F32 of one domain");
  if (ws < 1 || ws > kAdjMaxWs || !(ws & 1) || K < 1 || K > kAdjMaxK)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk_loss_adjoint: ws %d (odd, <= %d), K %d (<= %d)", ws, kAdjMaxWs,
                     K, kAdjMaxK);
  if (dL_dI1 && (dL_dI1->dtype != CUIMG_F32 || dL_dI1->nrows != I1->nrows || dL_dI1->ncols != I1->ncols))
    return set_error(ctx, CUIMG_E_INVALID, "lk_loss_adjoint: dL_dI1 must be F32 of the frame domain");
  const int nr = I1->nrows, nc = I1->ncols;
  const size_t npx = (size_t)nr * nc;
  if (npx >= (size_t)1 << 31) return set_error(ctx, CUIMG_E_UNSUPPORTED, "lk_loss_adjoint: frame too large");
  if (scratch_floats < 2 * npx)
    return set_error(ctx, CUIMG_E_INVALID, "lk_loss_adjoint: scratch holds %zu floats, needs %zu", scratch_floats,
                     2 * npx);
  adj_args a{};
  a.I0 = static_cast<const float*>(I0->begin);
  a.I1 = static_cast<const float*>(I1->begin);
  a.pitch0 = I0->pitch;
  a.pitch1 = I1->pitch;
  a.nr = nr;
  a.nc = nc;
  a.pos = reinterpret_cast<const float2*>(pos2);
  a.d = reinterpret_cast<const float2*>(d2);
  a.n = n;
  a.ws = ws;
  a.K = K;
  float coef[2 * kAdjMaxK + 1];
  CUIMG_TRY(gaussian_coefs(sigma, K, coef));
  for (int i = 0; i < 2 * K + 1; i++) a.c[i] = coef[i];
  a.S = scratch;
  a.loss_p = loss_p;
  a.dL_dd = reinterpret_cast<float2*>(dL_dd2);
  a.loss_total = loss_total;
  const int zb = (int)std::min<size_t>((npx + 255) / 256, 4096);
  CUIMG_TRY(launch(ctx, "adjoint_zero", [&] {
    zero_kernel<<<zb, 256, 0, ctx->stream>>>(scratch, npx);
    zero_kernel<<<1, 64, 0, ctx->stream>>>(reinterpret_cast<float*>(loss_total), 2);
  }));
  const int blocks = (int)std::min<uint32_t>((n + 3) / 4, 4096);
  if (n) CUIMG_TRY(launch(ctx, "lk_adjoint", [&] { lk_adjoint_kernel<<<blocks, 256, 0, ctx->stream>>>(a); }));
  if (dL_dI1) {
    float* tmp = scratch + npx;
    CUIMG_TRY(launch(ctx, "blur_adjoint", [&] {
      blur_adjoint_pass<<<zb, 256, 0, ctx->stream>>>(scratch, nc * 4, tmp, nc * 4, nr, nc, 0, K, a);
      blur_adjoint_pass<<<zb, 256, 0, ctx->stream>>>(tmp, nc * 4, static_cast<float*>(dL_dI1->begin),
                                                     dL_dI1->pitch, nr, nc, 1, K, a);
    }));
  }
  return 0;
}

}  // namespace cuimg

using namespace cuimg;

extern "C" int cuimg_hip_lk_loss_adjoint(cuimg_hip_ctx* ctx, const cuimg_img* I0, const cuimg_img* I1,
                                         const float* pos2, const float* d2, uint32_t n, int sigma, int half_width,
                                         int ws, double* loss_total, float* loss_p, float* dL_dd2,
                                         const cuimg_img* dL_dI1, float* scratch, size_t scratch_floats) {
  return lk_loss_adjoint(ctx, I0, I1, pos2, d2, n, sigma, half_width, ws, loss_total, loss_p, dL_dd2, dL_dI1,
                         scratch, scratch_floats);
}
