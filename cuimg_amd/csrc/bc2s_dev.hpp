// bc2s strategy device code for gfx950: the feature images (contrast, two
// integer 1-2-1 blurs), the bc2s descriptor, the two-step gradient-descent
// matcher, the per-cell flow statistics and their fusion over the scales, the
// bad-particle filter and the integer-tap Gaussian.  All of it is integer
// work (the flow statistics carry integer speeds in float), bit-exact with the
// reference's CPU strategy.
//
// Reference semantics (tracking2/ unless noted):
//  * bc2s_feature<cpu>::update built without OpenCV (bc2s_feature.hpp:232-253):
//    s1 = blur121(in), s2 = blur121(s1), then both borders clamp-filled; the
//    blur (cpu/gaussian_blur.h:9-47) is a row pass then a column pass of
//    (2 v + left + right) / 4 in int.  The reference leaves the first / last
//    column of the row pass and the first / last row of the column pass
//    unwritten; here they follow the same formula with the missing neighbour
//    replaced by the edge pixel (clamp-to-edge), as the oracle does;
//  * contrast_kernel (tracking_strategies.hpp:104-136): the sum of |a - b| over
//    the 4 opposite pixel pairs at distance 2; stored here saturated to 8 bits,
//    which keeps both of its uses exact (contrast < 3 in the matcher, < 5 in
//    the detector mask);
//  * compute_feature / distance (bc2s_feature.hpp:33-143): 8 samples of s1 on
//    every other point of the radius-3 ring (circle_r3_h[0, 2, .., 14]) and 8
//    of s2 at twice the offsets; every descriptor the strategy builds has all
//    its weights at 255, so the weighted distance is the plain sum of absolute
//    differences (s1 ring + s2 ring at scale 1, s2 ring only otherwise);
//  * gradient_descent_match / two_step_gradient_descent_match
//    (gradient_descent_matcher.hpp:9-125) with c8_h and c8_it_h
//    (neighb2d_data.h:80-85, 162-180);
//  * match_particles_kernel (tracking_strategies.hpp:274-402), flow statistics
//    and fusion (:537-634, 688-701), filter_bad_particles_kernel (:491-525).
#pragma once

#include "particles_dev.hpp"
#include "tracker_state.hpp"

namespace cuimg {

// Neighbour tables as packed immediates (a per-lane index into a constant
// array would cost a memory round trip before every candidate):
//  c8_h row / col + 1 in 2-bit fields, c8_it_h first / end in 4-bit fields
//  (index 8 = the first scan, over all 8 neighbours).
__device__ __forceinline__ int c8r(int i) { return (int)((0x9224u >> (2 * i)) & 3u) - 1; }  // -1,0,1,-1,1,-1,0,1
__device__ __forceinline__ int c8c(int i) { return (int)((0x016Au >> (2 * i)) & 3u) - 1; }  // 1,1,1,0,0,-1,-1,-1
__device__ __forceinline__ int c8it_a(int k) { return (int)((0x064422006ull >> (4 * k)) & 15u); }
__device__ __forceinline__ int c8it_b(int k) { return (int)((0x011775533ull >> (4 * k)) & 15u); }
// every other point of the radius-3 ring (compile-time, for unrolled loops)
constexpr int kB2RingR[8] = {-3, -2, 0, 2, 3, 2, 0, -2};
constexpr int kB2RingC[8] = {0, 2, 3, 2, 0, -2, -3, -2};

constexpr float kNoFlowV = -500.f;  // NO_FLOW (tracker_defines.h:4)
constexpr int kCell = 8;            // generic_strategy::flow_ratio

// ------------------------------------------------------------ feature images
// One output tile of kB2W x kB2H pixels per 256-thread workgroup.  Every
// stage works on clamped coordinates (a value at (r, c) outside the domain is
// the one at the nearest domain pixel), which is what both blurs' edge rule
// and the clamp-filled borders give; the input is read at clamped
// coordinates, so its own border is not used.
constexpr int kB2W = 64, kB2H = 32;

struct bc2s_prep_args {
  const uint8_t* in;  // u8 level
  int in_pitch;
  uint8_t* con;       // saturated contrast (u8) or nullptr
  int con_pitch;
  uint32_t* con32;    // full contrast (u32) or nullptr
  int con32_pitch;
  uint8_t* s1;
  int s1_pitch, s1_border;  // the border (<= 3) is clamp-filled
  uint8_t* s2;              // nullptr: s1 only (one blur)
  int s2_pitch, s2_border;  // (<= 6)
  uint4* desc;              // per-pixel descriptor (compute_feature at every pixel) or nullptr
  int desc_pitch;           // in uint4
  int nr, nc, tiles_x;
};

// Writes v at (r, c) and, where (r, c) lies on the domain edge, at the border
// pixels that replicate it.
__device__ __forceinline__ void put_clamped(uint8_t* img, int pitch, int border, int nr, int nc, int r, int c,
                                            uint8_t v) {
  const int ra = r == 0 ? -border : r, rb = r == nr - 1 ? nr - 1 + border : r;
  const int ca = c == 0 ? -border : c, cb = c == nc - 1 ? nc - 1 + border : c;
  for (int y = ra; y <= rb; y++)
    for (int x = ca; x <= cb; x++) *px<uint8_t>(img, pitch, y, x) = v;
}


// Stages (rows x columns around the tile, all at clamped coordinates):
//   tin  input          H+16 x W+16   (halo 8)
//   tt   row pass of in  H+16 x W+14
//   s1e  s1              H+14 x W+14   (halo 7: the descriptor reads s1 at +-3,
//   t2   row pass of s1  H+14 x W+12    s2 at +-6, s2 needs s1 at +-1)
//   s2e  s2              H+12 x W+12   (halo 6)
__device__ inline void bc2s_prep_tile(const bc2s_prep_args& a, int tile) {
  constexpr int W = kB2W, H = kB2H;
  constexpr int IW = W + 16, IH = H + 16;
  constexpr int TW = W + 14;
  constexpr int EW = W + 14, EH = H + 14;
  constexpr int UW = W + 12;
  constexpr int SW = W + 12, SH = H + 12;
  __shared__ uint8_t tin[IH * IW];
  __shared__ uint8_t tt[IH * TW];
  __shared__ uint8_t s1e[EH * EW];
  __shared__ uint8_t t2[EH * UW];
  __shared__ uint8_t s2e[SH * SW];
  const int nr = a.nr, nc = a.nc;
  const int r0 = (tile / a.tiles_x) * H, c0 = (tile % a.tiles_x) * W;
  const bool want_s2 = a.s2 || a.desc;
  for (int i = threadIdx.x; i < IH * IW; i += blockDim.x) {
    const int r = clampi(r0 - 8 + i / IW, 0, nr - 1), c = clampi(c0 - 8 + i % IW, 0, nc - 1);
    tin[i] = *px<uint8_t>(a.in, a.in_pitch, r, c);
  }
  __syncthreads();
  // row pass of in at clamped columns: tt(row, c0-7+j) = t(row, clamp(c0-7+j))
  for (int i = threadIdx.x; i < IH * TW; i += blockDim.x) {
    const int lr = i / TW;
    const int C = clampi(c0 - 7 + i % TW, 0, nc - 1);
    const int Cl = max(C - 1, 0), Cr = min(C + 1, nc - 1);
    const uint8_t* row = &tin[lr * IW - (c0 - 8)];
    tt[i] = (uint8_t)((2 * row[C] + row[Cl] + row[Cr]) >> 2);
  }
  __syncthreads();
  // column pass: s1e(r, c) = s1(clamp r, clamp c)
  for (int i = threadIdx.x; i < EH * EW; i += blockDim.x) {
    const int R = clampi(r0 - 7 + i / EW, 0, nr - 1), lc = i % EW;
    const int Ru = max(R - 1, 0), Rd = min(R + 1, nr - 1);
    auto T = [&](int y) { return (int)tt[(y - (r0 - 8)) * TW + lc]; };
    s1e[i] = (uint8_t)((2 * T(R) + T(Ru) + T(Rd)) >> 2);
  }
  __syncthreads();
  if (want_s2) {
    // row pass of s1: t2(r, c0-6+j) = t2(clamp r, clamp(c0-6+j))
    for (int i = threadIdx.x; i < EH * UW; i += blockDim.x) {
      const int lr = i / UW;
      const int C = clampi(c0 - 6 + i % UW, 0, nc - 1);
      const int Cl = max(C - 1, 0), Cr = min(C + 1, nc - 1);
      const uint8_t* row = &s1e[lr * EW - (c0 - 7)];
      t2[i] = (uint8_t)((2 * row[C] + row[Cl] + row[Cr]) >> 2);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SH * SW; i += blockDim.x) {
      const int R = clampi(r0 - 6 + i / SW, 0, nr - 1), lc = i % SW;
      const int Ru = max(R - 1, 0), Rd = min(R + 1, nr - 1);
      auto T = [&](int y) { return (int)t2[(y - (r0 - 7)) * UW + lc]; };
      s2e[i] = (uint8_t)((2 * T(R) + T(Ru) + T(Rd)) >> 2);
    }
    __syncthreads();
  }
  auto S1 = [&](int r, int c) { return (uint32_t)s1e[(r - (r0 - 7)) * EW + (c - (c0 - 7))]; };
  auto S2 = [&](int r, int c) { return (uint32_t)s2e[(r - (r0 - 6)) * SW + (c - (c0 - 6))]; };
  for (int i = threadIdx.x; i < H * W; i += blockDim.x) {
    const int r = r0 + i / W, c = c0 + i % W;
    if (r >= nr || c >= nc) continue;
    put_clamped(a.s1, a.s1_pitch, a.s1_border, nr, nc, r, c, (uint8_t)S1(r, c));
    if (a.s2) put_clamped(a.s2, a.s2_pitch, a.s2_border, nr, nc, r, c, (uint8_t)S2(r, c));
    if (a.desc) {
      uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < 8; k++) {
        w[k >> 2] |= S1(r + kB2RingR[k], c + kB2RingC[k]) << (8 * (k & 3));
        w[2 + (k >> 2)] |= S2(r + 2 * kB2RingR[k], c + 2 * kB2RingC[k]) << (8 * (k & 3));
      }
      a.desc[(size_t)r * a.desc_pitch + c] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    if (a.con || a.con32) {
      auto V = [&](int y, int x) {
        return (int)tin[(clampi(y, 0, nr - 1) - (r0 - 8)) * IW + (clampi(x, 0, nc - 1) - (c0 - 8))];
      };
      const int ct = abs(V(r, c + 2) - V(r, c - 2)) + abs(V(r - 2, c) - V(r + 2, c)) +
                     abs(V(r - 2, c - 2) - V(r + 2, c + 2)) + abs(V(r - 2, c + 2) - V(r + 2, c - 2));
      if (a.con) *px<uint8_t>(a.con, a.con_pitch, r, c) = (uint8_t)min(ct, 255);
      if (a.con32) *px<uint32_t>(a.con32, a.con32_pitch, r, c) = (uint32_t)ct;
    }
  }
}

// ------------------------------------------------------------ descriptor
// bc2s_internals::compute_feature (tracking2/bc2s_feature.hpp:119-143): the
// s1 samples at every other point of the radius-3 ring (circle_r3_h[0, 2, ..,
// 14], cuimg/neighb2d_data.h:184-189, (row, col)) and the s2 samples at twice
// the offsets; bytes 0..7 from s1, 8..15 from s2, packed little-endian into 4
// words.
__device__ __forceinline__ uint4 bc2s_feature_at(const bc2s_view& F, int r, int c) {
  constexpr int R8[8] = {-3, -2, 0, 2, 3, 2, 0, -2}, C8[8] = {0, 2, 3, 2, 0, -2, -3, -2};
  const uint8_t* p1 = px<uint8_t>(F.s1, F.p1, r, c);
  const uint8_t* p2 = px<uint8_t>(F.s2, F.p2, r, c);
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t a = p1[R8[i] * F.p1 + C8[i]];
    const uint32_t b = p2[2 * R8[i] * F.p2 + 2 * C8[i]];
    w[i >> 2] |= a << (8 * (i & 3));
    w[2 + (i >> 2)] |= b << (8 * (i & 3));
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}


// Sum of |ring sample - descriptor byte|: the s2 ring, plus the s1 ring at
// scale 1 (v_sad_u8 over packed bytes).
__device__ __forceinline__ int bc2s_distance_at(const bc2s_view& F, const uint4& f, int r, int c, bool scale1) {
  uint32_t q[4] = {0, 0, 0, 0};
  const uint8_t* p1 = px<uint8_t>(F.s1, F.p1, r, c);
  const uint8_t* p2 = px<uint8_t>(F.s2, F.p2, r, c);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    q[2 + (i >> 2)] |= (uint32_t)p2[2 * kB2RingR[i] * F.p2 + 2 * kB2RingC[i]] << (8 * (i & 3));
    if (scale1) q[i >> 2] |= (uint32_t)p1[kB2RingR[i] * F.p1 + kB2RingC[i]] << (8 * (i & 3));
  }
  uint32_t d = __builtin_amdgcn_sad_u8(q[2], f.z, 0u);
  d = __builtin_amdgcn_sad_u8(q[3], f.w, d);
  if (scale1) {
    d = __builtin_amdgcn_sad_u8(q[0], f.x, d);
    d = __builtin_amdgcn_sad_u8(q[1], f.y, d);
  }
  return (int)d;
}

struct descent_out {
  int dr, dc;
  float dist;
};

// gradient_descent_match: from (pr, pc), at most 10 moves to the strictly best
// in-domain c8 neighbour, the first scan over all 8, later ones over the
// c8_it_h side of the last improving neighbour; the starting distance is the
// scale-1 one.  Outside the domain: the prediction itself, 999999.
__device__ inline descent_out bc2s_descent(const bc2s_view& F, const uint4& f, int pr0, int pc0, bool scale1) {
  const int nr = F.nr, nc = F.nc;
  auto has = [&](int r, int c) { return r >= 0 && r < nr && c >= 0 && c < nc; };
  if (!has(pr0, pc0)) return descent_out{pr0, pc0, 999999.f};
  int pr = pr0, pc = pc0, mr = pr0, mc = pc0, mi = 8;
  int md = bc2s_distance_at(F, f, pr0, pc0, true);
  for (int search = 0; search < 10; search++) {
    int i = c8it_a(mi);
    const int end = c8it_b(mi);
    // the first neighbour is scanned before the end test (c8_it_h[8] = {0, 0})
    do {
      const int r = pr + c8r(i), c = pc + c8c(i);
      if (has(r, c)) {
        const int d = bc2s_distance_at(F, f, r, c, scale1);
        if (d < md) {
          mr = r;
          mc = c;
          mi = i;
          md = d;
        }
      }
      i = (i + 1) & 7;
    } while (i != end);
    if (pr == mr && pc == mc) break;
    pr = mr;
    pc = mc;
  }
  return descent_out{mr - pr0, mc - pc0, (float)md};
}

__device__ inline descent_out bc2s_two_step(const bc2s_view& F, const uint4& f, int pr, int pc) {
  const descent_out m1 = bc2s_descent(F, f, pr, pc, false);
  const descent_out m2 = bc2s_descent(F, f, pr + m1.dr, pc + m1.dc, true);
  return descent_out{m1.dr + m2.dr, m1.dc + m2.dc, m2.dist};
}

// The same descents with 8 lanes per particle: lane k scores the k-th
// candidate of a scan, and the lanes take the smallest (distance, scan
// position), which is the candidate the sequential strict-less scan keeps.
// One memory round trip per move instead of one per candidate.
__device__ __forceinline__ uint32_t grp8_min(uint32_t v) {
#pragma unroll
  for (int m = 1; m < 8; m <<= 1) v = min(v, (uint32_t)__shfl_xor((int)v, m, 8));
  return v;
}

__device__ inline descent_out bc2s_descent8(const bc2s_view& F, const uint4& f, int pr0, int pc0, bool scale1,
                                            int k) {
  int pr = pr0, pc = pc0, mi = 8;
  int md = bc2s_distance_at(F, f, pr0, pc0, true);
  for (int search = 0; search < 10; search++) {
    const int start = c8it_a(mi);
    const int len = mi == 8 ? 8 : ((c8it_b(mi) - start) & 7);
    uint32_t key = 0xFFFFFFFFu;
    if (k < len) {
      const int ci = (start + k) & 7;
      const int r = pr + c8r(ci), c = pc + c8c(ci);
      if (r >= 0 && r < F.nr && c >= 0 && c < F.nc)
        key = ((uint32_t)bc2s_distance_at(F, f, r, c, scale1) << 3) | (uint32_t)k;
    }
    key = grp8_min(key);
    if (key == 0xFFFFFFFFu || (int)(key >> 3) >= md) break;
    const int ci = (start + (int)(key & 7u)) & 7;
    pr += c8r(ci);
    pc += c8c(ci);
    mi = ci;
    md = (int)(key >> 3);
  }
  return descent_out{pr - pr0, pc - pc0, (float)md};
}

// The same 8-lane descent on the per-pixel descriptor image: a candidate's
// distance is the byte SAD of its descriptor (s2 half at scale 2).
__device__ __forceinline__ int sad_desc(const uint4& q, const uint4& f, bool scale1) {
  uint32_t d = __builtin_amdgcn_sad_u8(q.z, f.z, 0u);
  d = __builtin_amdgcn_sad_u8(q.w, f.w, d);
  if (scale1) {
    d = __builtin_amdgcn_sad_u8(q.x, f.x, d);
    d = __builtin_amdgcn_sad_u8(q.y, f.y, d);
  }
  return (int)d;
}

__device__ __forceinline__ uint4 desc_at(const uint4* D, int dp, int r, int c, bool scale1) {
  const uint4* q = D + (size_t)r * dp + c;
  if (scale1) return *q;
  const uint2 h = *(reinterpret_cast<const uint2*>(q) + 1);
  return make_uint4(0u, 0u, h.x, h.y);
}

__device__ inline descent_out bc2s_descent8_desc(const uint4* D, int dp, int nr, int nc, const uint4& f, int pr0,
                                                 int pc0, bool scale1, int k) {
  int pr = pr0, pc = pc0, mi = 8;
  int md = sad_desc(D[(size_t)pr0 * dp + pc0], f, true);
  for (int search = 0; search < 10; search++) {
    const int start = c8it_a(mi);
    const int len = mi == 8 ? 8 : ((c8it_b(mi) - start) & 7);
    uint32_t key = 0xFFFFFFFFu;
    if (k < len) {
      const int ci = (start + k) & 7;
      const int r = pr + c8r(ci), c = pc + c8c(ci);
      if (r >= 0 && r < nr && c >= 0 && c < nc)
        key = ((uint32_t)sad_desc(desc_at(D, dp, r, c, scale1), f, scale1) << 3) | (uint32_t)k;
    }
    key = grp8_min(key);
    if (key == 0xFFFFFFFFu || (int)(key >> 3) >= md) break;
    const int ci = (start + (int)(key & 7u)) & 7;
    pr += c8r(ci);
    pc += c8c(ci);
    mi = ci;
    md = (int)(key >> 3);
  }
  return descent_out{pr - pr0, pc - pc0, (float)md};
}

// ------------------------------------------------------------ frame step
// Per-scale job of the bc2s strategy: images of the current frame, the
// per-cell statistics, and the particle vectors with their descriptors.
struct bc2s_job {
  bc2s_prep_args prep[2];  // [frame slot]
  bc2s_view F;
  const uint4* desc;       // per-pixel descriptors of the current frame
  int desc_pitch;          // in uint4
  const uint8_t* con;
  int con_pitch;
  uint4* feat[2];       // [vector], descriptors parallel to the particles
  int cr, cc;           // cells (domain_div_up(domain, 8))
  uint4* stats;         // per cell: count, speed (row, col) bits, 0
  float2* flow;         // per cell
  uint32_t* msc;        // per cell: multiscale count
  int k;                // n_diff threshold (set_k)
  int last;             // coarsest scale of the stream (flow_fusion_kernel_root)
  int32_t* dec_accept;  // per particle decisions of the last match
  float2* dec_tr;
  float* dec_dist;
};

// match_particles_kernel, 8 lanes per particle: particles i = first,
// first + stride, ... < n (group indices) are moved / removed in place, the
// sparse image keeps the largest index of a pixel (the last writer in index
// order), decisions for the caller.  Every lane of a group follows the same
// path; lane 0 writes.
__device__ inline void bc2s_match_body(const bc2s_job& B, cuimg_particle_f* parts, uint4* feat, uint32_t n,
                                       uint32_t* sparse, int sparse_pitch, int32_t* dec_accept, float2* dec_tr,
                                       float* dec_dist, uint32_t first, uint32_t stride, uint32_t* live,
                                       uint32_t* accepted) {
  const bc2s_view F = B.F;
  const int k = (int)(threadIdx.x & 7u);
  for (uint32_t i = first; i < n; i += stride) {
    cuimg_particle_f p = parts[i];
    const uint4 f = feat[i];  // issued with the particle load
    int accept = 0;
    float2 tr = make_float2(0.f, 0.f);
    float dist = 0.f;
    if (p.age > 0) {
      const int r = (int)p.pos_r, c = (int)p.pos_c;
      const int pr = r + (p.age == 1 ? 0 : (int)p.speed_r), pc = c + (p.age == 1 ? 0 : (int)p.speed_c);
      if (pr >= 0 && pr < F.nr && pc >= 0 && pc < F.nc) {
        // two_step_gradient_descent_match: s2 ring, then both rings
        const descent_out m1 = bc2s_descent8_desc(B.desc, B.desc_pitch, F.nr, F.nc, f, pr, pc, false, k);
        const descent_out m2 = bc2s_descent8_desc(B.desc, B.desc_pitch, F.nr, F.nc, f, pr + m1.dr, pc + m1.dc, true, k);
        const int dr = m1.dr + m2.dr, dc = m1.dc + m2.dc;
        const int mr = pr + dr, mc = pc + dc;
        dist = m2.dist;
        tr = make_float2((float)(pr - r + dr), (float)(pc - c + dc));
        if (mr >= 0 && mr < F.nr && mc >= 0 && mc < F.nc) {
          const uint4 mf = B.desc[(size_t)mr * B.desc_pitch + mc];
          int n_diff = 0;
          const uint32_t a[2] = {f.z, f.w}, b[2] = {mf.z, mf.w};
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int x = (int)((a[j >> 2] >> (8 * (j & 3))) & 0xFFu), y = (int)((b[j >> 2] >> (8 * (j & 3))) & 0xFFu);
            n_diff += abs(x - y) > B.k;
          }
          if (n_diff < 4 && *px<uint8_t>(B.con, B.con_pitch, mr, mc) >= 3) {
            accept = 1;
            p.age++;
            p.speed_r = tr.x;
            p.speed_c = tr.y;
            p.pos_r = (float)mr;
            p.pos_c = (float)mc;
            if (k == 0) {
              if (tr.x != 0.f || tr.y != 0.f) feat[i] = mf;
              atomicMax(reinterpret_cast<int*>(px<uint32_t>(sparse, sparse_pitch, mr, mc)), (int)i);
            }
          }
        }
      }
      if (!accept) p.age = 0;
      if (k == 0) {
        parts[i] = p;
        ++*live;
        *accepted += accept;
      }
    }
    if (k == 0) {
      dec_accept[i] = accept;
      dec_tr[i] = tr;
      dec_dist[i] = dist;
    }
  }
}

// compute_flow_stats_kernel: the number of pixels of each 8x8 cell whose
// sparse entry is a particle of age >= 1, and the speed of the last of them in
// raster order.  One 256-thread workgroup per strip of one cell row and 256
// columns (32 cells): thread t reads column c0 + t of the 8 rows (coalesced
// rows of the sparse image), then the 8 threads of a cell combine their counts
// and their last hits (rank r * 8 + column in the cell).
__device__ inline void bc2s_stats_strip(const bc2s_job& B, const cuimg_particle_f* parts, const uint32_t* sparse,
                                        int sparse_pitch, int strip) {
  const int strips_x = (B.F.nc + 255) / 256;
  const int R = strip / strips_x, c = (strip % strips_x) * 256 + (int)threadIdx.x;
  uint32_t count = 0, rank = 0;  // rank + 1 of the last hit, 0 = none
  float sr = 0.f, sc = 0.f;
  if (c < B.F.nc) {
    const int rmax = min(kCell, B.F.nr - R * kCell);
    for (int r = 0; r < rmax; r++) {
      const uint32_t s = *px<uint32_t>(sparse, sparse_pitch, R * kCell + r, c);
      if (s == kEmpty) continue;
      const cuimg_particle_f q = parts[s];
      if (q.age >= 1) {
        count++;
        rank = (uint32_t)(r * kCell + (c & (kCell - 1))) + 1u;
        sr = q.speed_r;
        sc = q.speed_c;
      }
    }
  }
#pragma unroll
  for (int m = 1; m < 8; m <<= 1) {
    const uint32_t oc = (uint32_t)__shfl_xor((int)count, m, 8), orank = (uint32_t)__shfl_xor((int)rank, m, 8);
    const float osr = __shfl_xor(sr, m, 8), osc = __shfl_xor(sc, m, 8);
    count += oc;
    if (orank > rank) {
      rank = orank;
      sr = osr;
      sc = osc;
    }
  }
  if ((threadIdx.x & 7u) == 0 && c < B.F.nc) {
    const int C = c / kCell;
    B.stats[(size_t)R * B.cc + C] =
        make_uint4(count, __float_as_uint(rank ? sr : 0.f), __float_as_uint(rank ? sc : 0.f), 0u);
  }
}

__device__ __forceinline__ bool is_no_flow(float2 v) { return v.x == kNoFlowV && v.y == kNoFlowV; }

// flow_fusion_kernel / flow_fusion_kernel_root for the cells of scale s, from
// the statistics of scales s .. last of the stream (jobs js .. js + last - s):
// the flow of a cell is its own last speed when it counts more than one
// particle, else twice the coarser cell's flow when that one has a flow; the
// coarsest scale's is the last speed / count (count > 1).  An upper cell
// outside the coarser grid reads as empty.
template <class JobAt>
__device__ inline void bc2s_fuse_body(const bc2s_job& B, int s, JobAt job_at, uint32_t first, uint32_t stride) {
  const uint32_t ncell = (uint32_t)B.cr * B.cc;
  for (uint32_t x = first; x < ncell; x += stride) {
    const int R = (int)(x / (uint32_t)B.cc), C = (int)(x % (uint32_t)B.cc);
    const uint4 st = B.stats[x];
    // multiscale count
    uint32_t m = st.x;
    if (s < B.last) {
      const bc2s_job& U = job_at(s + 1);
      if (R / 2 < U.cr && C / 2 < U.cc) m += U.stats[(size_t)(R / 2) * U.cc + C / 2].x;
    }
    B.msc[x] = m;
    // flow: walk up to the scale that decides it, then apply the x2 steps down
    int t = s, y = R, z = C;
    float2 v = make_float2(kNoFlowV, kNoFlowV);
    for (;;) {
      const bc2s_job& T = job_at(t);
      const uint4 q = T.stats[(size_t)y * T.cc + z];
      if (t == B.last) {
        if (q.x > 1)
          v = make_float2(__fdiv_rn(__uint_as_float(q.y), (float)q.x), __fdiv_rn(__uint_as_float(q.z), (float)q.x));
        break;
      }
      if (q.x > 1) {
        v = make_float2(__uint_as_float(q.y), __uint_as_float(q.z));
        break;
      }
      const bc2s_job& U = job_at(t + 1);
      if (y / 2 >= U.cr || z / 2 >= U.cc) break;
      y /= 2;
      z /= 2;
      t++;
    }
    for (; t > s; t--) v = is_no_flow(v) ? v : make_float2(v.x * 2.f, v.y * 2.f);
    B.flow[x] = v;
  }
}

// filter_bad_particles_kernel on particles first, first + stride, ... < n.
__device__ inline void bc2s_filter_body(const bc2s_job& B, cuimg_particle_f* parts, uint32_t n, uint32_t first,
                                        uint32_t stride) {
  for (uint32_t i = first; i < n; i += stride) {
    cuimg_particle_f p = parts[i];
    const int R = (int)p.pos_r / kCell, C = (int)p.pos_c / kCell;
    const uint32_t m = B.msc[(size_t)R * B.cc + C];
    bool kill = false;
    if (p.age > 2 && m > 0) {
      const float2 f = B.flow[(size_t)R * B.cc + C];
      kill = norml2_rn(__fsub_rn(p.speed_r, f.x), __fsub_rn(p.speed_c, f.y)) > 10.f;
    }
    if (!kill && m == 1) {
      bool alone = true;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int y = R + c8r(k), z = C + c8c(k);
        if (y >= 0 && y < B.cr && z >= 0 && z < B.cc && B.msc[(size_t)y * B.cc + z] != 0) alone = false;
      }
      kill = alone;
    }
    if (kill && p.age != 0) {
      p.age = 0;
      parts[i].age = 0;
    }
  }
}

// ------------------------------------------------------------ integer Gaussian
// gaussian_kernel<cuda_gpu> taps applied separably (columns, then rows), each
// pass divided by the tap sum with truncation into u8, clamp-to-edge reads.
constexpr int kGiMaxHw = 15;
struct gint_args {
  const uint8_t* in;
  int in_pitch;
  uint8_t* tmp;
  int tmp_pitch;
  uint8_t* out;
  int out_pitch;
  int nr, nc, hw, sum;
  int w[2 * kGiMaxHw + 1];
};

__device__ inline void gint_pass(const gint_args& a, bool columns, int r, int c) {
  const uint8_t* src = columns ? a.in : a.tmp;
  const int sp = columns ? a.in_pitch : a.tmp_pitch;
  int acc = 0;
  for (int k = -a.hw; k <= a.hw; k++) {
    const int y = columns ? clampi(r + k, 0, a.nr - 1) : r, x = columns ? c : clampi(c + k, 0, a.nc - 1);
    acc += a.w[k + a.hw] * (int)*px<uint8_t>(src, sp, y, x);
  }
  const uint8_t v = (uint8_t)min(255, acc / a.sum);
  if (columns) *px<uint8_t>(a.tmp, a.tmp_pitch, r, c) = v;
  else *px<uint8_t>(a.out, a.out_pitch, r, c) = v;
}

}  // namespace cuimg
