// Particle container device code for gfx950: detector mask, box
// non-maximum selection, raster-order insertion, stream compaction and the
// trajectory merge.  Bodies take their block index and block count, so the
// same code runs as a standalone launch or as one job of a batched launch.
//
// Reference semantics (tracking2/):
//  * create_detector_mask + particle_mask_kernel (pyrlk_tracker.hpp:879-888,
//    tracking_strategies.hpp:201-225): mask <- 255, then 0 on the 3x3 block of
//    every particle of the dense vector, dead ones included;
//  * fast_detector::new_particles(cpu) (fast_detector.hpp:217-262): one box per
//    cell of domain/box at p*box + (1,1); skipped if any pixel holds a particle;
//    first strict maximum in raster order; rejected below the float threshold;
//    particles appended in raster order over domain - border(3)
//    (particle_container::add, particle_container.hpp:370-387);
//  * particle_container::compact(cpu) (particle_container.hpp:216-252): keep
//    age != 0 in order, matches[old] = new (or -1), sparse(pos) = new index with
//    the last particle of a pixel winning;
//  * merge_trajectories (merge_trajectories.h:10-35): in ascending index order a
//    live particle removes the first c8 neighbour (CPU table order,
//    neighb2d_data.h:80-85) more than 2 frames older whose speed differs by
//    less than 2 px.
//
// Raster order on the GPU: a box row covers `box` image rows, so the raster
// order of new points is (box row, row inside the box, box column).  The
// workgroup that selects a box row also counts its points per (box row, row
// offset) "sub-row", a prefix over sub-rows gives the offsets, and one wave per
// sub-row writes its points with a ballot prefix.  Compaction is a count and a
// scatter over 1024-particle blocks (with a scan launch between them in the
// standalone entry points; the frame step sums the prefixes in the scatter).  The serial merge is resolved
// exactly in parallel rounds and a serial replay of the few actors they leave
// open.
#pragma once

#include "tracker_state.hpp"

#include <climits>

namespace cuimg {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ int tzf(float v) { return (int)v; }

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread over the workgroup (blockDim <= 1024).
__device__ inline uint32_t block_excl_scan(uint32_t x, uint32_t* tmp, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  uint32_t inc = wave_incl_scan(x);
  if (lane == 63) tmp[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < nw ? tmp[lane] : 0;
    uint32_t wi = wave_incl_scan(w);
    if (lane < nw) tmp[32 + lane] = wi - w;
    if (lane == nw - 1) tmp[31] = wi;
  }
  __syncthreads();
  uint32_t r = inc - x + tmp[32 + wave];
  if (total) *total = tmp[31];
  __syncthreads();
  return r;
}

// Sum of x[lo, hi) over the wave; every lane gets it.
__device__ inline uint32_t wave_range_sum(const uint32_t* x, int lo, int hi) {
  uint32_t s = 0;
  for (int i = lo + (int)(threadIdx.x & 63); i < hi; i += 64) s += x[i];
  return (uint32_t)__shfl(wave_incl_scan(s), 63, 64);
}

// ------------------------------------------------------------------ fills
// Fills one row segment [p, p + n) of 4-byte words with v: unaligned head and
// tail words one by one, the aligned body with 16-byte stores; workgroup
// `part` of `nparts` takes every nparts-th block of the body.
__device__ __forceinline__ void fill_row_u32(uint32_t* p, int n, uint32_t v, int part, int nparts) {
  const int head = min(n, (int)(((16 - ((uintptr_t)p & 15)) & 15) >> 2));
  const int body = (n - head) >> 2;
  const int tail = n - head - 4 * body;
  if (part == 0 && (int)threadIdx.x < head) p[threadIdx.x] = v;
  uint4* q = reinterpret_cast<uint4*>(p + head);
  const uint4 v4 = make_uint4(v, v, v, v);
  for (int j = part * blockDim.x + threadIdx.x; j < body; j += nparts * blockDim.x) q[j] = v4;
  if (part == 0 && (int)threadIdx.x < tail) p[head + 4 * body + threadIdx.x] = v;
}

// 32-bit fill of rows [r0, r1) x cols [c0, c1) (coordinates relative to
// begin; negative ones reach into the border): one row per workgroup at a
// time, or a single row split over all workgroups.
__device__ inline void fill_u32_body(uint32_t* begin, int pitch, int r0, int r1, int c0, int c1, uint32_t v,
                                     int bid, int nb) {
  if (r1 - r0 == 1) {
    fill_row_u32(px<uint32_t>(begin, pitch, r0, c0), c1 - c0, v, bid, nb);
    return;
  }
  for (int r = r0 + bid; r < r1; r += nb) fill_row_u32(px<uint32_t>(begin, pitch, r, c0), c1 - c0, v, 0, 1);
}

// u8 domain fill (rows 4-byte aligned at column 0 or not): head bytes, then
// 16-byte stores, then tail bytes, one row per workgroup at a time.
__device__ inline void fill_u8_body(uint8_t* begin, int pitch, int nr, int nc, uint32_t v4, int bid, int nb) {
  for (int r = bid; r < nr; r += nb) {
    uint8_t* p = px<uint8_t>(begin, pitch, r, 0);
    const int head = min(nc, (int)((16 - ((uintptr_t)p & 15)) & 15));
    const int body = (nc - head) >> 4;
    const int tail = nc - head - 16 * body;
    if ((int)threadIdx.x < head) p[threadIdx.x] = (uint8_t)v4;
    uint4* q = reinterpret_cast<uint4*>(p + head);
    const uint4 w = make_uint4(v4, v4, v4, v4);
    for (int j = threadIdx.x; j < body; j += blockDim.x) q[j] = w;
    if ((int)threadIdx.x < tail) p[head + 16 * body + threadIdx.x] = (uint8_t)v4;
  }
}

// ------------------------------------------------------------ detector mask
// Zeroes the 3x3 block around the truncated position of each of the n
// particles in a u8 image (a detector mask that holds 255, or a saliency).
__device__ inline void particle_mask_core(const cuimg_particle_f* parts, uint32_t n, uint8_t* mask, int pitch,
                                          int nr, int nc, int bid, int nb) {
  for (uint32_t i = bid * blockDim.x + threadIdx.x; i < n; i += nb * blockDim.x) {
    const int r = tzf(parts[i].pos_r), c = tzf(parts[i].pos_c);
    for (int dr = -1; dr <= 1; dr++)
      for (int dc = -1; dc <= 1; dc++) {
        const int rr = r + dr, cc = c + dc;
        if (rr >= 0 && rr < nr && cc >= 0 && cc < nc) *px<uint8_t>(mask, pitch, rr, cc) = 0;
      }
  }
}

// ---------------------------------------------------------- box selection
__device__ __forceinline__ uint32_t sparse_get(const uint32_t* sp, int pitch, int nr, int nc, int r, int c) {
  if (r < 0 || r >= nr || c < 0 || c >= nc) return kEmpty;
  return *px<uint32_t>(sp, pitch, r, c);
}

// Box at (pr, pc): code dr*box + dc of the selected point, 0xFF when the box
// yields nothing.  Pixels outside the domain read as empty (sparse) and 0
// (saliency); a 0 never beats a maximum, which starts at the box's first
// pixel (the saliency border is 0, reached only when box == 1).
__device__ __forceinline__ uint8_t nms_code(const detect_args& a, float fast_th, int pr, int pc, int B, float vmax,
                                            int best) {
  const int mr = pr + best / B, mc = pc + best % B;
  if (!(vmax < fast_th) && mr >= 3 && mr <= a.nr - 4 && mc >= 3 && mc <= a.nc - 4) return (uint8_t)best;
  return 0xFF;
}

// Box side B <= 6: every sparse and saliency read of the box is issued before
// the first comparison (one memory round trip per box).
template <int B>
__device__ __forceinline__ uint8_t nms_box_fixed(const detect_args& a, float fast_th, int pr, int pc) {
  uint32_t sp[B * B];
  uint32_t sv[B * B];
#pragma unroll
  for (int r = 0; r < B; r++)
#pragma unroll
    for (int c = 0; c < B; c++) {
      const bool in = pr + r < a.nr && pc + c < a.nc;
      sp[r * B + c] = in ? *px<uint32_t>(a.sparse, a.sparse_pitch, pr + r, pc + c) : kEmpty;
      sv[r * B + c] = in ? *px<uint8_t>(a.sal, a.sal_pitch, pr + r, pc + c) : 0u;
    }
  bool skip = false;
#pragma unroll
  for (int k = 0; k < B * B; k++) skip |= sp[k] != kEmpty;
  if (skip) return 0xFF;
  float vmax = (float)sv[0];
  int best = 0;
#pragma unroll
  for (int k = 1; k < B * B; k++) {
    const float v = (float)sv[k];
    if (vmax < v) {
      vmax = v;
      best = k;
    }
  }
  return nms_code(a, fast_th, pr, pc, B, vmax, best);
}

// Any box side: one row of the box per memory round trip.
__device__ inline uint8_t nms_box_rows(const detect_args& a, float fast_th, int pr, int pc, int B) {
  bool skip = false;
  float vmax = 0.f;
  int best = 0;
  for (int r = 0; r < B; r++) {
    uint32_t sp[kMaxBox], sv[kMaxBox];
#pragma unroll
    for (int c = 0; c < kMaxBox; c++) {
      const bool in = c < B && pr + r < a.nr && pc + c < a.nc;
      sp[c] = in ? *px<uint32_t>(a.sparse, a.sparse_pitch, pr + r, pc + c) : kEmpty;
      sv[c] = in ? *px<uint8_t>(a.sal, a.sal_pitch, pr + r, pc + c) : 0u;
    }
#pragma unroll
    for (int c = 0; c < kMaxBox; c++) {
      skip |= sp[c] != kEmpty;
      const float v = (float)sv[c];
      if (r == 0 && c == 0) {
        vmax = v;
      } else if (c < B && vmax < v) {
        vmax = v;
        best = r * B + c;
      }
    }
  }
  return skip ? (uint8_t)0xFF : nms_code(a, fast_th, pr, pc, B, vmax, best);
}

// Box side 3 (the tracker's default): a thread takes four neighbouring boxes,
// columns 12q + 1 .. 12q + 12 of the box row, and reads the 16 columns
// 12q .. 12q + 15 of each of its 3 rows with one 16-byte load of saliency and
// four of the sparse image (pixel (0,0) and the pitch of both 16-byte aligned,
// checked by the caller).  Quads
// whose 16 columns do not all lie inside the domain take the per-box path.
// Rows are read up to nr - 1 (row nr, reached when 3 nbr = nr, is masked).
// Pixels outside the domain read as empty and 0, as in nms_box_fixed.  Codes
// and sub-row counts are those of the per-box path.
__device__ inline void nms_row3_body(const detect_args& a, float fast_th, int bi, uint32_t* cnts) {
  const int nq = (a.nbc + 3) >> 2;
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    const int pr = bi * 3 + 1, c0 = q * 12;
    if (c0 + 16 > a.nc) {
      for (int bj = 4 * q; bj < min(4 * q + 4, a.nbc); bj++) {
        const uint8_t code = nms_box_fixed<3>(a, fast_th, pr, bj * 3 + 1);
        a.boxcode[(size_t)bi * a.nbc + bj] = code;
        if (code != 0xFF) atomicAdd(&cnts[code / 3], 1u);
      }
      continue;
    }
    uint32_t sp[3][16];
    uint32_t sv[3][4];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int rl = min(pr + r, a.nr - 1);  // a row past the domain is masked below
      const uint4* s4 = reinterpret_cast<const uint4*>(px<uint32_t>(a.sparse, a.sparse_pitch, rl, c0));
      const uint4 v = *reinterpret_cast<const uint4*>(px<uint8_t>(a.sal, a.sal_pitch, rl, c0));
      sv[r][0] = v.x;
      sv[r][1] = v.y;
      sv[r][2] = v.z;
      sv[r][3] = v.w;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint4 w = s4[k];
        sp[r][4 * k] = w.x;
        sp[r][4 * k + 1] = w.y;
        sp[r][4 * k + 2] = w.z;
        sp[r][4 * k + 3] = w.w;
      }
    }
    uint8_t codes[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int pc = c0 + 1 + 3 * b;
      bool skip = false;
      float vmax = 0.f;
      int best = 0;
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const int j = 1 + 3 * b + c;  // column within the 16 loaded
          const bool in = pr + r < a.nr && pc + c < a.nc;
          skip |= in && sp[r][j] != kEmpty;
          const float v = in ? (float)((sv[r][j >> 2] >> (8 * (j & 3))) & 0xFFu) : 0.f;
          if (r == 0 && c == 0) {
            vmax = v;
          } else if (vmax < v) {
            vmax = v;
            best = r * 3 + c;
          }
        }
      codes[b] = skip ? (uint8_t)0xFF : nms_code(a, fast_th, pr, pc, 3, vmax, best);
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int bj = 4 * q + b;
      if (bj < a.nbc) {
        a.boxcode[(size_t)bi * a.nbc + bj] = codes[b];
        if (codes[b] != 0xFF) atomicAdd(&cnts[codes[b] / 3], 1u);
      }
    }
  }
}

// One workgroup per box row bi; each thread one box at a time.  The row's
// points are counted per sub-row (row offset dr) into subcnt.
__device__ inline void nms_row_body(const detect_args& a, float fast_th, int bi) {
  __shared__ uint32_t cnts[kMaxBox];
  const int B = a.box;
  if (threadIdx.x < B) cnts[threadIdx.x] = 0;
  __syncthreads();
  const bool aligned16 = (((uintptr_t)a.sparse | (uintptr_t)a.sal | (uintptr_t)a.sparse_pitch |
                            (uintptr_t)a.sal_pitch) & 15) == 0;
  if (B == 3 && aligned16) {
    nms_row3_body(a, fast_th, bi, cnts);
    __syncthreads();
    if (threadIdx.x < B) a.subcnt[bi * B + threadIdx.x] = cnts[threadIdx.x];
    return;
  }
  for (int bj = threadIdx.x; bj < a.nbc; bj += blockDim.x) {
    const int pr = bi * B + 1, pc = bj * B + 1;
    uint8_t code;
    switch (B) {
      case 1: code = nms_box_fixed<1>(a, fast_th, pr, pc); break;
      case 2: code = nms_box_fixed<2>(a, fast_th, pr, pc); break;
      case 3: code = nms_box_fixed<3>(a, fast_th, pr, pc); break;
      case 4: code = nms_box_fixed<4>(a, fast_th, pr, pc); break;
      case 5: code = nms_box_fixed<5>(a, fast_th, pr, pc); break;
      case 6: code = nms_box_fixed<6>(a, fast_th, pr, pc); break;
      default: code = nms_box_rows(a, fast_th, pr, pc, B); break;
    }
    a.boxcode[(size_t)bi * a.nbc + bj] = code;
    if (code != 0xFF) atomicAdd(&cnts[code / B], 1u);
  }
  __syncthreads();
  if (threadIdx.x < B) a.subcnt[bi * B + threadIdx.x] = cnts[threadIdx.x];
}

// Compaction count: flags (alive, owner of its sparse pixel) and alive count
// per 1024-particle block (256 threads; thread t takes particles t, t + 256,
// t + 512, t + 768 of the block, so a wave's loads are contiguous).  Ownership
// is read before any index is rewritten.
__device__ inline void compact_count_body(const detect_args& a, int bid) {
  __shared__ uint32_t tmp[64];
  const uint32_t n = a.cnt->n;
  const uint32_t base = bid * kCompactBlock + threadIdx.x;
  uint32_t c = 0;
  for (int k = 0; k < 4; k++) {
    const uint32_t i = base + k * (kCompactBlock / 4);
    if (i < n) {
      const cuimg_particle_f p = a.parts[i];
      uint8_t f = 0;
      if (p.age != 0) {
        f = 1;
        const uint32_t s = sparse_get(a.sparse, a.sparse_pitch, a.nr, a.nc, tzf(p.pos_r), tzf(p.pos_c));
        if (s == i) f |= 2;
        c++;
      }
      a.flags[i] = f;
    }
  }
  uint32_t total;
  block_excl_scan(c, tmp, &total);
  if (threadIdx.x == 0) a.blockcnt[bid] = total;
}

// One workgroup: exclusive scans of the block counts and sub-row counts, then
// the container counters.
__device__ inline void compact_scan_body(const detect_args& a) {
  __shared__ uint32_t tmp[64];
  __shared__ uint32_t carry;
  uint32_t sums[2];
  for (int pass = 0; pass < 2; pass++) {
    uint32_t* arr = pass == 0 ? a.blockcnt : a.subcnt;
    const int len = pass == 0 ? a.nblocks : a.nbr * a.box;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < len; b0 += blockDim.x) {
      const int i = b0 + threadIdx.x;
      const uint32_t x = i < len ? arr[i] : 0;
      uint32_t total;
      const uint32_t ex = block_excl_scan(x, tmp, &total);
      if (i < len) arr[i] = carry + ex;
      __syncthreads();
      if (threadIdx.x == 0) carry += total;
      __syncthreads();
    }
    sums[pass] = carry;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    level_counters& C = *a.cnt;
    const uint32_t n_old = C.n, n_live = sums[0], n_new = sums[1];
    C.n_old = n_old;
    C.n_live = n_live;
    C.n_new = n_new;
    uint32_t total = n_live + n_new;
    if (total > a.cap) {
      C.overflow = 1;
      total = a.cap;
    }
    C.n = total;
  }
}

// Compaction scatter of 1024-particle block bid whose first kept particle goes
// to index off0: four rounds of 256 consecutive particles (thread t takes
// particle t of the round, so a wave's loads and stores are contiguous), each
// an exclusive scan of the kept flags after the previous rounds' total.
__device__ inline void compact_scatter_at(const detect_args& a, int bid, uint32_t off0, uint32_t n_old) {
  __shared__ uint32_t tmp[64];
  uint32_t off = off0;
  for (int k = 0; k < 4; k++) {
    const uint32_t i = bid * kCompactBlock + k * (kCompactBlock / 4) + threadIdx.x;
    const uint8_t f = i < n_old ? a.flags[i] : 0;
    uint32_t total;
    const uint32_t o = off + block_excl_scan(f & 1u, tmp, &total);
    if (i < n_old) {
      if (f & 1) {
        const cuimg_particle_f p = a.parts[i];
        a.parts_out[o] = p;
        if (a.feat) a.feat_out[o] = a.feat[i];
        a.matches[i] = o;
        if (f & 2) *px<uint32_t>(a.sparse, a.sparse_pitch, tzf(p.pos_r), tzf(p.pos_c)) = o;
      } else {
        a.matches[i] = kEmpty;
      }
    }
    off += total;
  }
}

// After compact_scan_body: blockcnt holds exclusive offsets.
__device__ inline void compact_scatter_body(const detect_args& a, int bid) {
  compact_scatter_at(a, bid, a.blockcnt[bid], a.cnt->n_old);
}

// The wave of sub-row `sub` visits its selected points in column order:
// emit(k, r, c) with k the point's rank among all new points, starting at k0.
template <class Emit>
__device__ inline void for_new_points_from(const detect_args& a, int sub, uint32_t k, Emit emit) {
  const int lane = threadIdx.x & 63;
  const int bi = sub / a.box, dr = sub % a.box;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int b0 = 0; b0 < a.nbc; b0 += 64) {
    const int bj = b0 + lane;
    const uint8_t code = bj < a.nbc ? a.boxcode[(size_t)bi * a.nbc + bj] : 0xFF;
    const bool hit = code != 0xFF && code / a.box == dr;
    const uint64_t m = __ballot(hit);
    if (hit) emit(k + (uint32_t)__popcll(m & below), bi * a.box + 1 + dr, bj * a.box + 1 + code % a.box);
    k += (uint32_t)__popcll(m);
  }
}

// One wave per sub-row (4 per 256-thread block); after compact_scan_body
// subcnt holds the scanned sub-row offsets.
template <class Emit>
__device__ inline void for_new_points(const detect_args& a, int bid, Emit emit) {
  const int sub = bid * 4 + (threadIdx.x >> 6);
  if (sub >= a.nbr * a.box) return;
  for_new_points_from(a, sub, a.subcnt[sub], emit);
}

// Frame-step compaction without a scan launch: blockcnt and subcnt hold raw
// counts, each block or wave sums the prefix it needs, and block 0 writes the
// container counters (nothing else in this launch reads them; the count launch
// saved the previous size in n_old).
__device__ inline void compact_scatter_step(const detect_args& a, int bid) {
  const uint32_t off0 = wave_range_sum(a.blockcnt, 0, bid);
  compact_scatter_at(a, bid, off0, a.cnt->n_old);
  if (bid == 0) {
    const uint32_t n_live = wave_range_sum(a.blockcnt, 0, a.nblocks);
    const uint32_t n_new = wave_range_sum(a.subcnt, 0, a.nbr * a.box);
    if (threadIdx.x == 0) {
      level_counters& C = *a.cnt;
      uint32_t total = n_live + n_new;
      C.n_live = n_live;
      C.n_new = n_new;
      if (total > a.cap) {
        C.overflow = 1;
        total = a.cap;
      }
      C.n = total;
    }
  }
}

// Appends the new points after the live particles of the compacted vector.
__device__ inline void insert_new_step(const detect_args& a, int bid) {
  const int sub = bid * 4 + (threadIdx.x >> 6);
  if (sub >= a.nbr * a.box) return;
  const uint32_t n_live = wave_range_sum(a.blockcnt, 0, a.nblocks);
  const uint32_t n_new = wave_range_sum(a.subcnt, 0, a.nbr * a.box);
  const uint32_t k0 = wave_range_sum(a.subcnt, 0, sub);
  const uint32_t n = min(n_live + n_new, a.cap), n_old = a.cnt->n_old;
  for_new_points_from(a, sub, k0, [&](uint32_t kk, int r, int c) {
    const uint32_t j = n_live + kk;
    if (j < n) {
      cuimg_particle_f p;
      p.speed_r = 0.f;
      p.speed_c = 0.f;
      p.pos_r = (float)r;
      p.pos_c = (float)c;
      p.age = 1;
      p.pad = 0;
      a.parts_out[j] = p;
      if (a.feat) a.feat_out[j] = a.desc[(size_t)r * a.desc_pitch + c];
      *px<uint32_t>(a.sparse, a.sparse_pitch, r, c) = j;
      a.matches[n_old + kk] = j;
    } else {
      a.matches[n_old + kk] = kEmpty;
    }
  });
}

// Standalone selection output: (row, col) of the first cap new points.
__device__ inline void select_points_body(const detect_args& a, int32_t* pts, uint32_t cap, int bid) {
  for_new_points(a, bid, [&](uint32_t kk, int r, int c) {
    if (kk < cap) {
      pts[2 * (size_t)kk] = r;
      pts[2 * (size_t)kk + 1] = c;
    }
  });
}

// ------------------------------------------------------------------ merge
//
// Serial semantics, resolved in parallel rounds.  With the ages before the
// merge, the edge i -> b ("i may remove b") is static: b is a c8 neighbour
// (through the sparse image) with age_b > age_i + 2 and |speed_i - speed_b| < 2.
// In index order, i acts iff nobody earlier removed it, and then removes the
// first of its targets that nobody earlier removed.  Only actors targeting a
// particle can remove it.  A round resolves every open actor i whose outcome
// no open actor j < i can change: no open j < i targets i, and no open j < i
// targets one of i's targets up to the first one still alive at i's turn
// (zeroer[b] = index of b's remover; b is dead at i's turn iff zeroer[b] < i).
// The smallest open actor always resolves, and on tracking data each round
// leaves about a quarter of the open actors.  Once at most 64 remain (or after
// kMaxRounds rounds) the rest are sorted by index and replayed serially by one
// wave: a removal is visible to a later actor through zeroer (removals by
// resolved actors and by earlier chunks) or through a ballot over the victims
// of the earlier lanes of its chunk.
//
// minunres[b] = (255 - round) << 24 | smallest open actor targeting b, written
// by atomicMin: a later round's tag is smaller, so stale entries never win and
// no reset pass is needed.  Particle indices stay below 2^24.  Between calls
// zeroer and minunres hold 0xFFFFFFFF; the resolve kernel restores the entries
// it touched.

constexpr uint32_t kNoIdx = 0xFFFFFFFFu;
constexpr uint32_t kPendLds = 4096;
constexpr uint32_t kMaxRounds = 250;
constexpr uint32_t kSerialSwitch = 64;

__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t round_tag(uint32_t r) { return (255u - r) << 24; }

// norml2 (cuimg/builtin_math.h:123-130): both squares rounded, then added.
__device__ __forceinline__ float norml2_rn(float a, float b) {
#pragma clang fp contract(off)
  return sqrtf(a * a + b * b);
}

__host__ __device__ inline uint32_t pow2_at_least(uint32_t n) {
  uint32_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// One thread per particle: targets in c8 order, actor records appended with
// one atomic per wave, round-1 minunres by atomicMin.
__device__ inline void merge_scan_body(const merge_args& a, int bid, int nb) {
  const uint32_t n = a.cnt->n;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = nb * blockDim.x;
  for (uint32_t i0 = bid * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
    const uint32_t i = i0 + lane;
    merge_actor rec;
    rec.i = (int32_t)i;
    rec.mask = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) rec.b[k] = -1;
    if (i < n) {
      const cuimg_particle_f p = a.parts[i];
      if (p.age != 0) {
        // c8 neighbours in the CPU table order (c8_h, neighb2d_data.h:80-85);
        // all eight sparse reads are issued first (clamped addresses, pixels
        // outside the domain read as empty), then the neighbours' particles
        constexpr int dr8[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
        constexpr int dc8[8] = {1, 1, 1, 0, 0, -1, -1, -1};
        const int r = tzf(p.pos_r), c = tzf(p.pos_c);
        uint32_t bs[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int rr = r + dr8[k], cc = c + dc8[k];
          const bool in = rr >= 0 && rr < a.nr && cc >= 0 && cc < a.nc;
          const uint32_t v = *px<uint32_t>(a.sparse, a.sparse_pitch, clampi(rr, 0, a.nr - 1), clampi(cc, 0, a.nc - 1));
          bs[k] = in && v < n ? v : kEmpty;  // indices past the vector are treated as empty
        }
        uint32_t qa[8];
        float qr[8], qc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          qa[k] = 0;
          qr[k] = qc[k] = 0.f;
          if (bs[k] != kEmpty) {
            const cuimg_particle_f q = a.parts[bs[k]];
            qa[k] = q.age;
            qr[k] = q.speed_r;
            qc[k] = q.speed_c;
          }
        }
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (bs[k] != kEmpty && (int)qa[k] > (int)p.age + 2 &&
              norml2_rn(__fsub_rn(p.speed_r, qr[k]), __fsub_rn(p.speed_c, qc[k])) < 2.f) {
            rec.mask |= 1u << k;
            rec.b[k] = (int32_t)bs[k];
          }
      }
    }
    const bool act = rec.mask != 0;
    const uint64_t m = __ballot(act);
    if (m == 0) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&a.cnt->actors, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (act) {
      const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
      a.actors[base + __popcll(m & below)] = rec;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (rec.mask & (1u << k)) atomicMin(&a.minunres[rec.b[k]], round_tag(1) | i);
    }
  }
}

// Round-r decision for an open actor: false = still open.  A concurrent
// resolver j of this round writes zeroer[b] only where no open actor below j
// targets b, so whatever this thread reads there gives the same answer.
__device__ inline bool merge_decide(const merge_args& a, const merge_actor& A, uint32_t r) {
  const uint32_t tg = 255u - r, i = (uint32_t)A.i;
  // every value the decision can need, read in one round trip (minunres does
  // not change during a round's decisions, and zeroer only where the answer
  // does not depend on it)
  const uint32_t mi = ld32(&a.minunres[i]), zi = ld32(&a.zeroer[i]);
  uint32_t mb[8], zb[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    mb[j] = zb[j] = kNoIdx;
    if (A.mask & (1u << j)) {
      mb[j] = ld32(&a.minunres[A.b[j]]);
      zb[j] = ld32(&a.zeroer[A.b[j]]);
    }
  }
  auto blocked = [&](uint32_t v) { return (v >> 24) == tg && (v & 0xFFFFFFu) < i; };
  if (blocked(mi)) return false;
  if (zi < i) return true;  // removed before its turn
  int32_t victim = -1;
#pragma unroll
  for (int j = 0; j < 8; j++)
    if (victim < 0 && (A.mask & (1u << j))) {
      if (blocked(mb[j])) return false;
      if (zb[j] < i) continue;
      victim = A.b[j];
    }
  if (victim >= 0) {
    st32(&a.zeroer[victim], i);
    a.parts[victim].age = 0;
  }
  return true;
}

// Bitonic sort of keys[0, np) ascending, np a power of two, by one workgroup.
template <bool GLOBAL>
__device__ inline void bitonic_sort(unsigned long long* keys, uint32_t np) {
  for (uint32_t size = 2; size <= np; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t t = threadIdx.x; t < np / 2; t += blockDim.x) {
        const uint32_t lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        const bool asc = (lo & size) == 0;
        const unsigned long long x = GLOBAL ? ld64(&keys[lo]) : keys[lo];
        const unsigned long long y = GLOBAL ? ld64(&keys[hi]) : keys[hi];
        if ((x > y) == asc) {
          if (GLOBAL) {
            st64(&keys[lo], y);
            st64(&keys[hi], x);
          } else {
            keys[lo] = y;
            keys[hi] = x;
          }
        }
      }
      if (GLOBAL) __threadfence_block();  // one workgroup
      __syncthreads();
    }
}

// Ascending sort of one key per lane across the wave.
__device__ inline unsigned long long wave_sort(unsigned long long key) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const unsigned long long o = __shfl_xor(key, j, 64);
      const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
      key = keep_min ? (o < key ? o : key) : (o > key ? o : key);
    }
  return key;
}

__device__ __forceinline__ uint32_t rdl(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// Serial replay of up to 64 open actors in index order (one wave; `key` of
// lane l = the l-th actor, kNoIdx keys past the end).
__device__ inline void replay_chunk(const merge_args& a, unsigned long long key) {
  const int lane = threadIdx.x & 63;
  const bool valid = key != ~0ull;
  uint32_t ai = kNoIdx, zs = kNoIdx, amask = 0;
  uint32_t ab[8], zb[8];
#pragma unroll
  for (int j = 0; j < 8; j++) ab[j] = zb[j] = kNoIdx;
  if (valid) {
    const merge_actor A = a.actors[(uint32_t)key];
    ai = (uint32_t)A.i;
    amask = A.mask;
    zs = ld32(&a.zeroer[ai]);
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (amask & (1u << j)) {
        ab[j] = (uint32_t)A.b[j];
        zb[j] = ld32(&a.zeroer[ab[j]]);
      }
  }
  const uint32_t cnt = (uint32_t)__popcll(__ballot(valid));  // valid keys come first
  uint32_t victim = kNoIdx;
  for (uint32_t l = 0; l < cnt; l++) {
    const uint32_t i = rdl(ai, l);
    const uint64_t below = l == 0 ? 0ull : (~0ull >> (64 - l));
    if (rdl(zs, l) < i || (__ballot(victim == i) & below)) continue;  // removed before its turn
    const uint32_t msk = rdl(amask, l);
    bool done = false;
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (!done && (msk & (1u << j))) {
        const uint32_t b = rdl(ab[j], l);
        if (!(rdl(zb[j], l) < i || (__ballot(victim == b) & below))) {
          if ((uint32_t)lane == l) victim = b;
          done = true;
        }
      }
  }
  if (valid && victim != kNoIdx) {
    st32(&a.zeroer[victim], ai);
    a.parts[victim].age = 0;
  }
  __threadfence_block();
}

// Round 1 for every actor, over many workgroups (the decisions of one round
// do not depend on each other's order): open actors are appended to the
// pending list a.pend, counted in cnt->pend1.
__device__ inline void merge_round1_body(const merge_args& a, int bid, int nb) {
  const uint32_t m = a.cnt->actors;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = nb * blockDim.x;
  for (uint32_t k0 = bid * blockDim.x + (threadIdx.x & ~63u); k0 < m; k0 += stride) {
    const uint32_t k = k0 + lane;
    bool open = false;
    unsigned long long key = 0;
    if (k < m) {
      const merge_actor A = a.actors[k];
      open = !merge_decide(a, A, 1);
      key = ((unsigned long long)(uint32_t)A.i << 32) | k;
    }
    const uint64_t bal = __ballot(open);
    if (bal == 0) continue;
    const int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&a.cnt->pend1, (uint32_t)__popcll(bal));
    base = __shfl(base, leader);
    if (open) st64(&a.pend[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1))], key);
  }
}

// Resets the zeroer / minunres entries of actors [0, count) to kNoIdx (the
// state between merges), workgroup bid of nb.
__device__ inline void merge_restore_body(const merge_args& a, uint32_t count, int bid, int nb) {
  for (uint32_t k = bid * blockDim.x + threadIdx.x; k < count; k += nb * blockDim.x) {
    const merge_actor A = a.actors[k];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (A.mask & (1u << j)) {
        st32(&a.minunres[A.b[j]], kNoIdx);
        st32(&a.zeroer[A.b[j]], kNoIdx);
      }
  }
}

// One workgroup: later rounds over the actors round 1 left open, serial replay
// of the rest, restoration of the scratch (or, in a frame step, the count of
// entries for the next step's fill to restore).  The workgroup is the only
// reader and writer of the scratch while it runs, so its fences are
// workgroup-scope.
__device__ inline void merge_resolve_body(const merge_args& a) {
  __shared__ unsigned long long keys_lds[kPendLds];
  __shared__ unsigned long long small[64];
  __shared__ uint32_t nopen[2], nsmall;
  const uint32_t m = a.cnt->actors;
  const uint32_t np = ld32(&a.cnt->pend1);
  const bool in_lds = np <= kPendLds;
  if (in_lds)
    for (uint32_t k = threadIdx.x; k < np; k += blockDim.x) keys_lds[k] = ld64(&a.pend[k]);
  __syncthreads();
  unsigned long long* keys = in_lds ? keys_lds : a.pend;
  auto kget = [&](uint32_t s) { return in_lds ? keys_lds[s] : ld64(&a.pend[s]); };
  auto kset = [&](uint32_t s, unsigned long long v) {
    if (in_lds) keys_lds[s] = v;
    else st64(&a.pend[s], v);
  };

  uint32_t open = np, r = 1;
  while (open > kSerialSwitch && r < kMaxRounds) {
    r++;
    if (threadIdx.x == 0) nopen[r & 1] = 0;
    for (uint32_t s = threadIdx.x; s < np; s += blockDim.x) {
      const unsigned long long key = kget(s);
      if (key == ~0ull) continue;
      const merge_actor A = a.actors[(uint32_t)key];
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (A.mask & (1u << j)) atomicMin(&a.minunres[A.b[j]], round_tag(r) | (uint32_t)A.i);
    }
    __threadfence_block();
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < np; s += blockDim.x) {
      const unsigned long long key = kget(s);
      if (key == ~0ull) continue;
      if (merge_decide(a, a.actors[(uint32_t)key], r)) kset(s, ~0ull);
      else atomicAdd(&nopen[r & 1], 1u);
    }
    __threadfence_block();
    __syncthreads();
    open = nopen[r & 1];
  }

  if (open > 0 && open <= 64) {
    // gather the open keys into one wave, sort them there, replay
    if (threadIdx.x == 0) nsmall = 0;
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < np; s += blockDim.x) {
      const unsigned long long key = kget(s);
      if (key != ~0ull) small[atomicAdd(&nsmall, 1u)] = key;
    }
    __syncthreads();
    if (threadIdx.x < 64) replay_chunk(a, wave_sort(threadIdx.x < open ? small[threadIdx.x] : ~0ull));
  } else if (open > 0) {
    // after kMaxRounds: full sort (open keys first), replay in chunks of 64
    const uint32_t np2 = pow2_at_least(np);
    for (uint32_t k = np + threadIdx.x; k < np2; k += blockDim.x) kset(k, ~0ull);
    __threadfence_block();
    __syncthreads();
    if (np2 <= kPendLds) bitonic_sort<false>(keys, np2);
    else bitonic_sort<true>(keys, np2);
    if (threadIdx.x < 64)
      for (uint32_t c0 = 0; c0 < open; c0 += 64) {
        const uint32_t idx = c0 + threadIdx.x;
        replay_chunk(a, idx < open ? kget(idx) : ~0ull);
      }
  }
  __syncthreads();
  // restore the scratch for the next call, or leave it to the next frame step
  if (a.defer_restore) {
    if (threadIdx.x == 0) a.cnt->merge_restore = m;
  } else {
    merge_restore_body(a, m, 0, 1);
  }
  if (threadIdx.x == 0) {
    a.cnt->actors = 0;
    a.cnt->pend1 = 0;
  }
}

}  // namespace cuimg
