// Particle container for gfx950: detector mask, box non-maximum selection,
// raster-order insertion, stream compaction and the trajectory merge.
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
// order of new points is (box row, row inside the box, box column).  Each
// (box row, row offset) "sub-row" is counted by one wave, one scan over
// sub-rows gives the offsets, and one wave per sub-row writes its points with
// a ballot prefix.  Compaction is a classic count / scan / scatter over
// 1024-particle blocks.  The serial merge is resolved exactly in parallel rounds
// (see merge_round below).
#include "tracker_state.hpp"

#include <climits>

namespace cuimg {

__constant__ int8_t kC8R[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int8_t kC8C[8] = {1, 1, 1, 0, 0, -1, -1, -1};

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
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t* tmp, uint32_t* total) {
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

// ------------------------------------------------------------------ fills
__global__ void fill_u32_kernel(uint32_t* begin, int pitch, int r0, int r1, int c0, int c1, uint32_t v) {
  const int w = c1 - c0;
  const long total = (long)(r1 - r0) * w;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
    *px<uint32_t>(begin, pitch, r0 + (int)(i / w), c0 + (int)(i % w)) = v;
}

int launch_fill_u32(cuimg_hip_ctx* ctx, uint32_t* begin, int pitch, int r0, int r1, int c0, int c1, uint32_t v) {
  const long total = (long)(r1 - r0) * (c1 - c0);
  if (total <= 0) return 0;
  const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
  return launch(ctx, "fill_u32", [&] {
    fill_u32_kernel<<<blocks, 256, 0, ctx->stream>>>(begin, pitch, r0, r1, c0, c1, v);
  });
}

// u8 domain fill with 4-byte stores (rows are 4-byte aligned at column 0).
__global__ void fill_u8_kernel(uint8_t* begin, int pitch, int nr, int nc, uint32_t v4) {
  const int wpr = (nc + 3) / 4;
  const long total = (long)nr * wpr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const int r = (int)(i / wpr), w = (int)(i % wpr);
    uint8_t* p = px<uint8_t>(begin, pitch, r, 4 * w);
    if (4 * w + 4 <= nc) {
      *reinterpret_cast<uint32_t*>(p) = v4;
    } else {
      for (int k = 0; 4 * w + k < nc; k++) p[k] = (uint8_t)v4;
    }
  }
}

int launch_fill_u8(cuimg_hip_ctx* ctx, uint8_t* begin, int pitch, int nr, int nc, uint8_t v) {
  const long total = (long)nr * ((nc + 3) / 4);
  if (total <= 0) return 0;
  const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
  const uint32_t v4 = 0x01010101u * v;
  return launch(ctx, "fill_u8", [&] { fill_u8_kernel<<<blocks, 256, 0, ctx->stream>>>(begin, pitch, nr, nc, v4); });
}

// ------------------------------------------------------------ detector mask
__global__ void particle_mask_kernel(detect_args a) {
  const uint32_t n = a.cnt->n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int r = tzf(a.parts[i].pos_r), c = tzf(a.parts[i].pos_c);
    for (int dr = -1; dr <= 1; dr++)
      for (int dc = -1; dc <= 1; dc++) {
        const int rr = r + dr, cc = c + dc;
        if (rr >= 0 && rr < a.nr && cc >= 0 && cc < a.nc) *px<uint8_t>(a.mask, a.mask_pitch, rr, cc) = 0;
      }
  }
}

int launch_particle_mask(cuimg_hip_ctx* ctx, const detect_args& a) {
  CUIMG_TRY(launch_fill_u8(ctx, a.mask, a.mask_pitch, a.nr, a.nc, 255));
  const int blocks = (int)std::min<uint32_t>((a.cap + 255) / 256, 2048);
  return launch(ctx, "particle_mask", [&] { particle_mask_kernel<<<blocks, 256, 0, ctx->stream>>>(a); });
}

// ---------------------------------------------------------- box selection
__device__ __forceinline__ uint32_t sparse_get(const uint32_t* sp, int pitch, int nr, int nc, int r, int c) {
  if (r < 0 || r >= nr || c < 0 || c >= nc) return kEmpty;
  return *px<uint32_t>(sp, pitch, r, c);
}

// One workgroup per box row; each thread one box.  code = dr*box + dc of the
// selected point, 0xFF when the box yields nothing.
__global__ __launch_bounds__(256) void nms_kernel(detect_args a) {
  const int bi = blockIdx.x;
  const int B = a.box;
  for (int bj = threadIdx.x; bj < a.nbc; bj += blockDim.x) {
    const int pr = bi * B + 1, pc = bj * B + 1;
    uint8_t code = 0xFF;
    bool skip = false;
    for (int r = 0; r < B && !skip; r++)
      for (int c = 0; c < B; c++)
        if (pr + r < a.nr && pc + c < a.nc &&
            *px<uint32_t>(a.sparse, a.sparse_pitch, pr + r, pc + c) != kEmpty) {
          skip = true;
          break;
        }
    if (!skip) {
      // the saliency border is 0 (its memset), reached only when box == 1
      float vmax = (pr < a.nr && pc < a.nc) ? (float)*px<uint8_t>(a.sal, a.sal_pitch, pr, pc) : 0.f;
      int mr = pr, mc = pc;
      for (int r = 0; r < B; r++)
        for (int c = 0; c < B; c++)
          if (pr + r < a.nr && pc + c < a.nc) {
            const float sv = (float)*px<uint8_t>(a.sal, a.sal_pitch, pr + r, pc + c);
            if (vmax < sv) {
              vmax = sv;
              mr = pr + r;
              mc = pc + c;
            }
          }
      if (!(vmax < a.fast_th) && mr >= 3 && mr <= a.nr - 4 && mc >= 3 && mc <= a.nc - 4)
        code = (uint8_t)((mr - pr) * B + (mc - pc));
    }
    a.boxcode[(size_t)bi * a.nbc + bj] = code;
  }
}

// One wave per sub-row (box row bi, row offset dr): number of points.
__global__ __launch_bounds__(256) void subrow_count_kernel(detect_args a) {
  const int sub = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (sub >= a.nbr * a.box) return;
  const int bi = sub / a.box, dr = sub % a.box;
  uint32_t count = 0;
  for (int b0 = 0; b0 < a.nbc; b0 += 64) {
    const int bj = b0 + lane;
    const uint8_t code = bj < a.nbc ? a.boxcode[(size_t)bi * a.nbc + bj] : 0xFF;
    const bool hit = code != 0xFF && code / a.box == dr;
    count += __popcll(__ballot(hit));
  }
  if (lane == 0) a.subcnt[sub] = count;
}

// Compaction count: flags (alive, owner of its sparse pixel) and alive count
// per 1024-particle block.  Ownership is read before any index is rewritten.
__global__ __launch_bounds__(256) void compact_count_kernel(detect_args a) {
  __shared__ uint32_t tmp[64];
  const uint32_t n = a.cnt->n;
  const uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t c = 0;
  for (int k = 0; k < 4; k++) {
    const uint32_t i = base + k;
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
  if (threadIdx.x == 0) a.blockcnt[blockIdx.x] = total;
}

// Single workgroup: exclusive scans of the block counts and sub-row counts,
// then the container counters.
__global__ __launch_bounds__(1024) void compact_scan_kernel(detect_args a) {
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

__global__ __launch_bounds__(256) void compact_scatter_kernel(detect_args a) {
  __shared__ uint32_t tmp[64];
  const uint32_t n_old = a.cnt->n_old;
  const uint32_t base = blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t c = 0;
  uint8_t f[4];
  for (int k = 0; k < 4; k++) {
    const uint32_t i = base + k;
    f[k] = i < n_old ? a.flags[i] : 0;
    c += f[k] & 1;
  }
  uint32_t off = a.blockcnt[blockIdx.x] + block_excl_scan(c, tmp, nullptr);
  for (int k = 0; k < 4; k++) {
    const uint32_t i = base + k;
    if (i >= n_old) break;
    if (f[k] & 1) {
      const cuimg_particle_f p = a.parts[i];
      a.parts_out[off] = p;
      a.matches[i] = off;
      if (f[k] & 2) *px<uint32_t>(a.sparse, a.sparse_pitch, tzf(p.pos_r), tzf(p.pos_c)) = off;
      off++;
    } else {
      a.matches[i] = kEmpty;
    }
  }
}

// One wave per sub-row: appends its points in column order.
__global__ __launch_bounds__(256) void insert_new_kernel(detect_args a) {
  const int sub = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (sub >= a.nbr * a.box) return;
  const level_counters& C = *a.cnt;
  const uint32_t n_live = C.n_live, n_old = C.n_old, n = C.n;
  const int bi = sub / a.box, dr = sub % a.box;
  uint32_t k = a.subcnt[sub];  // rank among new points
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int b0 = 0; b0 < a.nbc; b0 += 64) {
    const int bj = b0 + lane;
    const uint8_t code = bj < a.nbc ? a.boxcode[(size_t)bi * a.nbc + bj] : 0xFF;
    const bool hit = code != 0xFF && code / a.box == dr;
    const uint64_t m = __ballot(hit);
    if (hit) {
      const uint32_t kk = k + (uint32_t)__popcll(m & below);
      const uint32_t j = n_live + kk;
      if (j < n) {
        const int r = bi * a.box + 1 + dr, c = bj * a.box + 1 + code % a.box;
        cuimg_particle_f p;
        p.speed_r = 0.f;
        p.speed_c = 0.f;
        p.pos_r = (float)r;
        p.pos_c = (float)c;
        p.age = 1;
        p.pad = 0;
        a.parts_out[j] = p;
        *px<uint32_t>(a.sparse, a.sparse_pitch, r, c) = j;
        a.matches[n_old + kk] = j;
      } else {
        a.matches[n_old + kk] = kEmpty;
      }
    }
    k += (uint32_t)__popcll(m);
  }
}

int launch_detect_and_compact(cuimg_hip_ctx* ctx, const detect_args& a) {
  const int nsub = a.nbr * a.box;
  if (a.nbr > 0 && a.nbc > 0) {
    CUIMG_TRY(launch(ctx, "nms_boxes", [&] { nms_kernel<<<a.nbr, 256, 0, ctx->stream>>>(a); }));
    CUIMG_TRY(launch(ctx, "subrow_count",
                     [&] { subrow_count_kernel<<<ceil_div(nsub, 4), 256, 0, ctx->stream>>>(a); }));
  }
  CUIMG_TRY(launch(ctx, "compact_count", [&] { compact_count_kernel<<<a.nblocks, 256, 0, ctx->stream>>>(a); }));
  CUIMG_TRY(launch(ctx, "compact_scan", [&] { compact_scan_kernel<<<1, 1024, 0, ctx->stream>>>(a); }));
  CUIMG_TRY(launch(ctx, "compact_scatter", [&] { compact_scatter_kernel<<<a.nblocks, 256, 0, ctx->stream>>>(a); }));
  if (nsub > 0)
    CUIMG_TRY(launch(ctx, "insert_new", [&] { insert_new_kernel<<<ceil_div(nsub, 4), 256, 0, ctx->stream>>>(a); }));
  return 0;
}

// ------------------------------------------------------------------ merge
//
// Serial semantics, resolved in parallel.  With the ages before the merge, the
// edge i -> b ("i may remove b") is static: b is a c8 neighbour (through the
// sparse image) with age_b > age_i + 2 and |speed_i - speed_b| < 2.  In index
// order, i acts iff nobody earlier removed it, and then removes the first of
// its targets that nobody earlier removed.  A round resolves every pending
// actor i whose outcome no longer depends on a pending actor j < i (j < i
// targeting i, or j < i targeting one of i's targets up to the first live
// one); the smallest pending actor always resolves, so the rounds terminate,
// and each decision equals the serial one.

__device__ __forceinline__ int32_t ld_acq(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float norml2_rn(float a, float b) {
  return sqrtf(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)));
}

__global__ __launch_bounds__(256) void merge_init_kernel(merge_args a) {
  const uint32_t n = a.cnt->n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    st_rel(&a.zeroer[i], INT_MAX);
    st_rel(&a.minunres[i], INT_MAX);
    const cuimg_particle_f p = a.parts[i];
    if (p.age == 0) continue;
    const int r = tzf(p.pos_r), c = tzf(p.pos_c);
    uint32_t mask = 0;
    for (int k = 0; k < 8; k++) {
      const uint32_t b = sparse_get(a.sparse, a.sparse_pitch, a.nr, a.nc, r + kC8R[k], c + kC8C[k]);
      if (b == kEmpty || b >= n) continue;  // indices past the vector are treated as empty
      const cuimg_particle_f q = a.parts[b];
      if ((int)q.age > (int)p.age + 2 &&
          norml2_rn(__fsub_rn(p.speed_r, q.speed_r), __fsub_rn(p.speed_c, q.speed_c)) < 2.f)
        mask |= 1u << k;
    }
    if (mask) {
      const uint32_t slot = atomicAdd(&a.cnt->actors, 1u);
      a.actors[slot] = i | (mask << 24);
    }
  }
}

__device__ __forceinline__ int32_t target_of(const merge_args& a, const cuimg_particle_f& p, int k) {
  return (int32_t)sparse_get(a.sparse, a.sparse_pitch, a.nr, a.nc, tzf(p.pos_r) + kC8R[k], tzf(p.pos_c) + kC8C[k]);
}

__global__ __launch_bounds__(1024) void merge_resolve_kernel(merge_args a) {
  __shared__ uint32_t pending;
  const uint32_t m = a.cnt->actors;
  for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) a.resolved[k] = 0;
  __syncthreads();
  for (;;) {
    // (a) clear the pending-minimum of pending actors and their targets
    for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
      if (a.resolved[k]) continue;
      const uint32_t e = a.actors[k];
      const int32_t i = (int32_t)(e & 0xFFFFFFu);
      const uint32_t mask = e >> 24;
      st_rel(&a.minunres[i], INT_MAX);
      const cuimg_particle_f p = a.parts[i];
      for (int t = 0; t < 8; t++)
        if (mask & (1u << t)) st_rel(&a.minunres[target_of(a, p, t)], INT_MAX);
    }
    __threadfence();
    __syncthreads();
    // (b) smallest pending actor targeting each particle
    for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
      if (a.resolved[k]) continue;
      const uint32_t e = a.actors[k];
      const int32_t i = (int32_t)(e & 0xFFFFFFu);
      const uint32_t mask = e >> 24;
      const cuimg_particle_f p = a.parts[i];
      for (int t = 0; t < 8; t++)
        if (mask & (1u << t)) atomicMin(&a.minunres[target_of(a, p, t)], i);
    }
    __threadfence();
    if (threadIdx.x == 0) pending = 0;
    __syncthreads();
    // (c) resolve what no smaller pending actor can change
    for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
      if (a.resolved[k]) continue;
      const uint32_t e = a.actors[k];
      const int32_t i = (int32_t)(e & 0xFFFFFFu);
      const uint32_t mask = e >> 24;
      bool done = false;
      if (ld_acq(&a.minunres[i]) < i) {
        done = false;  // an earlier pending actor may remove i
      } else if (ld_acq(&a.zeroer[i]) < i) {
        done = true;   // removed before its turn: no action
      } else {
        const cuimg_particle_f p = a.parts[i];
        bool blocked = false;
        int32_t victim = -1;
        for (int t = 0; t < 8 && victim < 0 && !blocked; t++) {
          if (!(mask & (1u << t))) continue;
          const int32_t b = target_of(a, p, t);
          if (ld_acq(&a.minunres[b]) < i) blocked = true;
          else if (ld_acq(&a.zeroer[b]) < i) continue;
          else victim = b;
        }
        if (!blocked) {
          if (victim >= 0) atomicMin(&a.zeroer[victim], i);
          done = true;
        }
      }
      if (done) a.resolved[k] = 1;
      else atomicAdd(&pending, 1u);
    }
    __threadfence();
    __syncthreads();
    if (pending == 0) break;
    __syncthreads();
  }
  if (threadIdx.x == 0) a.cnt->actors = 0;
}

__global__ __launch_bounds__(256) void merge_apply_kernel(merge_args a) {
  const uint32_t n = a.cnt->n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (ld_acq(&a.zeroer[i]) != INT_MAX) a.parts[i].age = 0;
}

int launch_merge(cuimg_hip_ctx* ctx, const merge_args& a) {
  const int blocks = (int)std::min<uint32_t>((a.cap + 255) / 256, 2048);
  CUIMG_TRY(launch(ctx, "merge_init", [&] { merge_init_kernel<<<blocks, 256, 0, ctx->stream>>>(a); }));
  CUIMG_TRY(launch(ctx, "merge_resolve", [&] { merge_resolve_kernel<<<1, 1024, 0, ctx->stream>>>(a); }));
  return launch(ctx, "merge_apply", [&] { merge_apply_kernel<<<blocks, 256, 0, ctx->stream>>>(a); });
}

}  // namespace cuimg
