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

// c8_it_h: first neighbour and end of the scan after a move to neighbour k (8: start)
static __constant__ int8_t kC8ItA[9] = {6, 0, 0, 2, 2, 4, 4, 6, 0};
static __constant__ int8_t kC8ItB[9] = {3, 3, 5, 5, 7, 7, 1, 1, 0};

constexpr float kNoFlowV = -500.f;  // NO_FLOW (tracker_defines.h:4)
constexpr int kCell = 8;            // generic_strategy::flow_ratio

// ------------------------------------------------------------ feature images
// One output tile of kB2W x kB2H pixels per 256-thread workgroup.
constexpr int kB2W = 64, kB2H = 16;

struct bc2s_prep_args {
  const uint8_t* in;  // u8 level; without CLAMP its border is >= 2 and clamp-filled
  int in_pitch;
  uint8_t* con;       // saturated contrast (u8) or nullptr
  int con_pitch;
  uint32_t* con32;    // full contrast (u32) or nullptr
  int con32_pitch;
  uint8_t* s1;
  int s1_pitch, s1_border;  // the border (<= 3 read by the descriptor) is clamp-filled
  uint8_t* s2;              // nullptr: s1 only (one blur)
  int s2_pitch, s2_border;  // (<= 6)
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

template <bool CLAMP>
__device__ inline void bc2s_prep_tile(const bc2s_prep_args& a, int tile) {
  constexpr int IW = kB2W + 4, IH = kB2H + 4;  // input with a halo of 2
  constexpr int TW = kB2W + 2;                 // row pass, columns c0-1 .. c0+W
  constexpr int EW = kB2W + 2, EH = kB2H + 2;  // s1 with a halo of 1
  __shared__ uint8_t tin[IH * IW];
  __shared__ uint8_t tt[IH * TW];
  __shared__ uint8_t s1e[EH * EW];
  const int nr = a.nr, nc = a.nc;
  const int r0 = (tile / a.tiles_x) * kB2H, c0 = (tile % a.tiles_x) * kB2W;
  // input rows r0-2 .. r0+H+1, cols c0-2 .. c0+W+1; reads past the domain
  // take the edge pixel (a clamp-filled border, or clamped coordinates)
  for (int i = threadIdx.x; i < IH * IW; i += blockDim.x) {
    int r = r0 - 2 + i / IW, c = c0 - 2 + i % IW;
    if (CLAMP) {
      r = clampi(r, 0, nr - 1);
      c = clampi(c, 0, nc - 1);
    } else {
      r = clampi(r, -2, nr + 1);
      c = clampi(c, -2, nc + 1);
    }
    tin[i] = *px<uint8_t>(a.in, a.in_pitch, r, c);
  }
  __syncthreads();
  auto IN = [&](int r, int c) { return (int)tin[(r - (r0 - 2)) * IW + (c - (c0 - 2))]; };
  // row pass t(R, C) for R = r0-2 .. r0+H+1, C = c0-1 .. c0+W (clamped C)
  for (int i = threadIdx.x; i < IH * TW; i += blockDim.x) {
    const int R = r0 - 2 + i / TW;
    const int C = clampi(c0 - 1 + i % TW, 0, nc - 1);
    const int Cl = max(C - 1, 0), Cr = min(C + 1, nc - 1);
    tt[i] = (uint8_t)((2 * IN(R, C) + IN(R, Cl) + IN(R, Cr)) >> 2);
  }
  __syncthreads();
  // column pass at clamped coordinates: s1e(r, c) = s1(clamp r, clamp c)
  for (int i = threadIdx.x; i < EH * EW; i += blockDim.x) {
    const int R = clampi(r0 - 1 + i / EW, 0, nr - 1);
    const int lc = i % EW;  // column c0-1+lc, whose clamped row pass is tt column lc
    const int Ru = max(R - 1, 0), Rd = min(R + 1, nr - 1);
    auto T = [&](int y) { return (int)tt[(y - (r0 - 2)) * TW + lc]; };
    s1e[i] = (uint8_t)((2 * T(R) + T(Ru) + T(Rd)) >> 2);
  }
  __syncthreads();
  auto S1 = [&](int r, int c) { return (int)s1e[(r - (r0 - 1)) * EW + (c - (c0 - 1))]; };
  for (int i = threadIdx.x; i < kB2H * kB2W; i += blockDim.x) {
    const int r = r0 + i / kB2W, c = c0 + i % kB2W;
    if (r >= nr || c >= nc) continue;
    // s2 = blur121(s1): s1e already holds the clamped s1 around the tile
    auto T2 = [&](int y) { return (2 * S1(y, c) + S1(y, c - 1) + S1(y, c + 1)) >> 2; };
    const uint8_t v2 = (uint8_t)((2 * T2(r) + T2(r - 1) + T2(r + 1)) >> 2);
    put_clamped(a.s1, a.s1_pitch, a.s1_border, nr, nc, r, c, (uint8_t)S1(r, c));
    if (a.s2) put_clamped(a.s2, a.s2_pitch, a.s2_border, nr, nc, r, c, v2);
    if (a.con || a.con32) {
      auto V = [&](int y, int x) {
        return IN(CLAMP ? clampi(y, 0, nr - 1) : clampi(y, -2, nr + 1), CLAMP ? clampi(x, 0, nc - 1) : clampi(x, -2, nc + 1));
      };
      const int ct = abs(V(r, c + 2) - V(r, c - 2)) + abs(V(r - 2, c) - V(r + 2, c)) +
                     abs(V(r - 2, c - 2) - V(r + 2, c + 2)) + abs(V(r - 2, c + 2) - V(r + 2, c - 2));
      if (a.con) *px<uint8_t>(a.con, a.con_pitch, r, c) = (uint8_t)min(ct, 255);
      if (a.con32) *px<uint32_t>(a.con32, a.con32_pitch, r, c) = (uint32_t)ct;
    }
  }
}

// ------------------------------------------------------------ descriptor
// (bc2s_view and bc2s_feature_at: tracker_state.hpp, particles_dev.hpp)

// Sum of |ring sample - descriptor byte|: the s2 ring, plus the s1 ring at
// scale 1 (v_sad_u8 over packed bytes).
__device__ __forceinline__ int bc2s_distance_at(const bc2s_view& F, const uint4& f, int r, int c, bool scale1) {
  uint32_t q[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 8; i++) {
    q[2 + (i >> 2)] |= (uint32_t)*px<uint8_t>(F.s2, F.p2, r + 2 * kB2R[i], c + 2 * kB2C[i]) << (8 * (i & 3));
    if (scale1) q[i >> 2] |= (uint32_t)*px<uint8_t>(F.s1, F.p1, r + kB2R[i], c + kB2C[i]) << (8 * (i & 3));
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
    int i = kC8ItA[mi];
    const int end = kC8ItB[mi];
    // the first neighbour is scanned before the end test (c8_it_h[8] = {0, 0})
    do {
      const int r = pr + kC8R[i], c = pc + kC8C[i];
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

// ------------------------------------------------------------ frame step
// Per-scale job of the bc2s strategy: images of the current frame, the
// per-cell statistics, and the particle vectors with their descriptors.
struct bc2s_job {
  bc2s_prep_args prep[2];  // [frame slot]
  bc2s_view F;
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

// match_particles_kernel on particles i = first, first + stride, ... < n:
// move / remove in place, the sparse image keeps the largest index of a pixel
// (the last writer in index order), decisions for the caller.
__device__ inline void bc2s_match_body(const bc2s_job& B, cuimg_particle_f* parts, uint4* feat, uint32_t n,
                                       uint32_t* sparse, int sparse_pitch, int32_t* dec_accept, float2* dec_tr,
                                       float* dec_dist, uint32_t first, uint32_t stride, uint32_t* live,
                                       uint32_t* accepted) {
  const bc2s_view F = B.F;
  for (uint32_t i = first; i < n; i += stride) {
    cuimg_particle_f p = parts[i];
    int accept = 0;
    float2 tr = make_float2(0.f, 0.f);
    float dist = 0.f;
    if (p.age > 0) {
      const int r = (int)p.pos_r, c = (int)p.pos_c;
      const int pr = r + (p.age == 1 ? 0 : (int)p.speed_r), pc = c + (p.age == 1 ? 0 : (int)p.speed_c);
      if (pr >= 0 && pr < F.nr && pc >= 0 && pc < F.nc) {
        const uint4 f = feat[i];
        const descent_out m = bc2s_two_step(F, f, pr, pc);
        const int mr = pr + m.dr, mc = pc + m.dc;
        dist = m.dist;
        tr = make_float2((float)(pr - r + m.dr), (float)(pc - c + m.dc));
        if (mr >= 0 && mr < F.nr && mc >= 0 && mc < F.nc) {
          const uint4 mf = bc2s_feature_at(F, mr, mc);
          int n_diff = 0;
          const uint32_t a[2] = {f.z, f.w}, b[2] = {mf.z, mf.w};
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int x = (int)((a[k >> 2] >> (8 * (k & 3))) & 0xFFu), y = (int)((b[k >> 2] >> (8 * (k & 3))) & 0xFFu);
            n_diff += abs(x - y) > B.k;
          }
          if (n_diff < 4 && *px<uint8_t>(B.con, B.con_pitch, mr, mc) >= 3) {
            accept = 1;
            p.age++;
            p.speed_r = tr.x;
            p.speed_c = tr.y;
            p.pos_r = (float)mr;
            p.pos_c = (float)mc;
            if (tr.x != 0.f || tr.y != 0.f) feat[i] = mf;
            atomicMax(reinterpret_cast<int*>(px<uint32_t>(sparse, sparse_pitch, mr, mc)), (int)i);
          }
        }
      }
      if (!accept) p.age = 0;
      parts[i] = p;
      ++*live;
      *accepted += accept;
    }
    dec_accept[i] = accept;
    dec_tr[i] = tr;
    dec_dist[i] = dist;
  }
}

// compute_flow_stats_kernel for cells first, first + stride, ...: the number of
// pixels of the 8x8 cell whose sparse entry is a particle of age >= 1, and the
// speed of the last of them in raster order.
__device__ inline void bc2s_stats_body(const bc2s_job& B, const cuimg_particle_f* parts, const uint32_t* sparse,
                                       int sparse_pitch, uint32_t first, uint32_t stride) {
  const uint32_t ncell = (uint32_t)B.cr * B.cc;
  for (uint32_t x = first; x < ncell; x += stride) {
    const int R = (int)(x / (uint32_t)B.cc), C = (int)(x % (uint32_t)B.cc);
    uint32_t count = 0;
    float2 last = make_float2(0.f, 0.f);
    const int rmax = min(kCell, B.F.nr - R * kCell), cmax = min(kCell, B.F.nc - C * kCell);
    for (int r = 0; r < rmax; r++) {
      const uint32_t* row = px<uint32_t>(sparse, sparse_pitch, R * kCell + r, C * kCell);
      for (int c = 0; c < cmax; c++) {
        const uint32_t s = row[c];
        if (s == kEmpty) continue;
        const cuimg_particle_f q = parts[s];
        if (q.age >= 1) {
          count++;
          last = make_float2(q.speed_r, q.speed_c);
        }
      }
    }
    B.stats[x] = make_uint4(count, __float_as_uint(last.x), __float_as_uint(last.y), 0u);
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
        const int y = R + kC8R[k], z = C + kC8C[k];
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
