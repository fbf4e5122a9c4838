// tracker<pyrlk> frame step on the GPU, for one video stream or a batch of
// independent streams: per-scale state, the batched step kernels, graph
// capture and the C ABI of trackers and batches.
//
// Reference call stack (tracking2/): tracker::run (tracker.hpp:73-91) builds the
// input pyramid (update_input / subsample_input, :53-69) and calls the
// strategy's update on every scale, coarsest first; pyrlk_cpu::update
// (pyrlk_tracker.hpp:831-875) matches the particles (frame > 0), runs the
// detector and inserts new particles (every detector_frequency frames, with a
// compaction), then merges trajectories (every filtering_frequency frames).
// Scales only interact through the images (the recursive prediction re-solves
// coarser scales itself), so all scales of all streams run side by side.
//
// Design: every (stream, scale) pair is a "job" with a device-resident
// descriptor (level_job) that holds the argument blocks of each kernel for
// every ping-pong combination (frame slot, current particle vector).  Which
// combination applies, and whether this frame matches / detects / merges, is
// derived on the device from the job's level_counters and frequencies, so a
// frame step is the same fixed sequence of 12 launches for any number of
// streams and any frame: it is captured once into a hipGraph and replayed.
// Each batched launch maps its blocks to jobs through a per-kind prefix array
// of block counts.  The per-stream input frame pointers are the only per-step
// upload (a small table, copied from a pinned ring before the launch).
//
// GPU layout per scale: two u8 pyramid slots (current / previous frame, border
// 3, clamp-filled), two FP32 blurred slots for the blurred-frame profile, the
// detector's mask and saliency, the u32 sparse index image (border 7), two
// dense particle vectors (compaction ping-pongs between them), the matches
// vector, and per-particle scratch for compaction and the merge.
#include "bc2s_dev.hpp"
#include "blur_grad_dev.hpp"
#include "fast_dev.hpp"
#include "gaussian_dev.hpp"
#include "lk_dev.hpp"
#include "particles_dev.hpp"
#include "pyramid_dev.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace cuimg {

struct level_res {
  int nr = 0, nc = 0;
  cuimg_img u8[2]{};
  cuimg_img blur[2]{};
  cuimg_img grad[2]{};  // Sobel/8 of blur[k] (blurred-frame profile)
  cuimg_img sal{}, sparse{};
  uint8_t* boxcode = nullptr;
  uint32_t* subcnt = nullptr;
  cuimg_particle_f* parts[2] = {nullptr, nullptr};
  uint32_t* matches = nullptr;
  uint8_t* flags = nullptr;
  uint32_t* blockcnt = nullptr;
  int32_t* dec_accept = nullptr;
  float2* dec_tr = nullptr;
  float* dec_dist = nullptr;
  merge_args merge{};  // merge scratch (one allocation)
  uint32_t cap = 0;
  int nblocks = 0, nbr = 0, nbc = 0;
  cuimg_scale_params sp{};
  // bc2s strategy: contrast (saturated u8), feature images, descriptors of
  // both particle vectors, per-cell statistics / flow / multiscale count
  cuimg_img con{}, s1{}, s2{};
  uint4* desc = nullptr;  // per-pixel descriptors (nc per row)
  uint4* feat[2] = {nullptr, nullptr};
  int cr = 0, cc = 0;
  uint4* stats = nullptr;
  float2* flow = nullptr;
  uint32_t* msc = nullptr;
};

struct job_params {
  int det_freq, filt_freq, with_merge, fast_n;
  float fast_th;
  int bc2s;  // generic_strategy with the bc2s descriptor (else pyrlk)
};

// Sobel/8 gradient of a blurred level `in` into `out` (its own pass or with the blur).
struct grad_job {
  const float* in;
  float2* out;
  int in_pitch, out_pitch;
  int nr, nc;
};

struct level_job {
  int stream, s;
  int nr, nc;
  job_params prm;
  level_counters* cnt;
  level_counters* cnt0;  // the stream's scale 0 (its step count sets the frame slot)
  uint8_t* u8[2];
  int u8_pitch;
  uint8_t* sal;
  int sal_pitch;
  uint32_t* sparse;
  int sparse_pitch, sparse_border;
  int fill_sparse_blocks;  // blocks of the sparse reset
  int fast_tiles_x;
  lk_args lk[4];        // [slot * 2 + cur]
  detect_args det[2];   // [cur]
  merge_args mg[2];     // [current vector after the detector]
  gauss_job blur[2];    // [slot]
  grad_job grad[2];     // [slot]
  bc2s_job bc;          // bc2s strategy
};

struct stream_job {
  int first, nlev;
  pyr_args pyr[2];  // [slot]; the frame is taken from the input table
};

struct frame_in {
  const void* begin;
  int pitch;
  int vec_ok;
};

enum {
  K_PYR, K_BLUR, K_FILL, K_LK, K_PMASK, K_FAST, K_DETC, K_SCATI, K_MSCAN, K_MR1,
  K_BPREP, K_BMATCH, K_FSTAT, K_FUSE, K_FILT, K_NKINDS
};
constexpr int kRing = 8;

}  // namespace cuimg

// Host-frame ingestion of one stream: frames go through two pinned staging
// buffers to two device frames on the batch's copy stream, so the upload of
// frame t+1 runs while the step of frame t does.  channels > 1: the packed
// interleaved frame is copied whole and channel 0 extracted on the device.
struct host_ingest {
  int channels = 0;
  uint8_t* pinned[2] = {nullptr, nullptr};
  void* raw[2] = {nullptr, nullptr};
  cuimg_img gray[2]{};
  hipEvent_t copied[2] = {nullptr, nullptr};    // the upload from pinned[k] is done
  hipEvent_t consumed[2] = {nullptr, nullptr};  // the step that read gray[k] is done
  int k = 0;
};

struct cuimg_hip_batch {
  cuimg_hip_ctx* ctx = nullptr;
  int nstreams = 0, njobs = 0;
  cuimg_tracker_params P0{};                 // matcher / blur settings shared by all streams
  std::vector<cuimg_tracker_params> P;       // per stream
  std::vector<int> first;                    // job index of each stream's scale 0
  std::vector<cuimg::level_res> lv;          // per job
  std::vector<cuimg::level_job> hj;          // host copy of the job table
  cuimg::level_job* dj = nullptr;
  std::vector<cuimg::stream_job> hs;
  cuimg::stream_job* ds = nullptr;
  cuimg::level_counters* cnt = nullptr;      // device, one per job
  cuimg::frame_in* din = nullptr;            // device input table
  cuimg::frame_in* hin = nullptr;            // pinned ring of input tables
  hipEvent_t ring_ev[cuimg::kRing]{};
  // FAST runs on a second stream (default priority) beside the matcher:
  // forked after the pyramid, joined before the saliency is masked
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  uint64_t steps = 0;
  std::vector<uint32_t> prefix_h[cuimg::K_NKINDS];
  uint32_t* prefix_d[cuimg::K_NKINDS]{};   // prefix + block -> job map per kind
  size_t table_cap[cuimg::K_NKINDS]{};
  uint32_t total[cuimg::K_NKINDS]{};
  cuimg::coef_pack cp{};
  int lk_ppw = 4;     // particles per matcher wave (tuning)
  int blur_th = 0;    // blur tile height, 0 = heuristic (tuning)
  bool use_graph = true;
  hipGraphExec_t exec = nullptr;        // frame step with the pyramid of the staged frame
  hipGraphExec_t exec_step = nullptr;   // frame step on the input built by update_input
  bool input_ready = false;             // update_input staged a frame the next step has not used
  std::vector<host_ingest> ingest;           // per stream, for run_host
  hipStream_t copy = nullptr;                // host-frame uploads (default priority)
  std::vector<void*> allocs;
  std::vector<cuimg_hip_tracker*> views;
};

struct cuimg_hip_tracker {
  cuimg_hip_batch* b = nullptr;
  int stream = 0;
  bool owner = false;
};

namespace cuimg {

// ------------------------------------------------------------ step kernels
// Block -> job map of a batched launch.  The per-kind table holds the
// prefix of block counts (njobs + 1 words) followed by one word per block:
// (job << kJobShift) | block index within the job, so a workgroup finds its
// job with one load instead of a binary search over the prefix.
constexpr int kJobShift = 20;
constexpr uint32_t kJobBlockMask = (1u << kJobShift) - 1;

struct job_block {
  int j, lb;
};

__device__ __forceinline__ job_block job_of_block(const uint32_t* table, int njobs) {
  const uint32_t e = uni(table[njobs + 1 + blockIdx.x]);
  return job_block{(int)(e >> kJobShift), (int)(e & kJobBlockMask)};
}

struct step_flags {
  uint32_t frame, cur, slot;
  bool match, detect, filter, merge;
};

// pyrlk_cpu::update matches from the second frame on; generic_strategy::update
// matches (and resets the sparse image) at every frame, and filters on the
// filtering frames.
__device__ __forceinline__ step_flags flags_of(const level_job& J) {
  step_flags f;
  const job_params prm = uni(J.prm);
  f.frame = uni(J.cnt->frame);
  f.cur = uni(J.cnt->cur) & 1u;
  f.slot = uni(J.cnt0->steps) & 1u;
  f.match = prm.bc2s || f.frame > 0;
  f.detect = f.frame % (uint32_t)prm.det_freq == 0;
  f.filter = prm.bc2s && f.frame % (uint32_t)prm.filt_freq == 0;
  f.merge = prm.with_merge && f.frame % (uint32_t)prm.filt_freq == 0;
  return f;
}

#define CUIMG_JOB(prefix, njobs)                               \
  const job_block jb_ = job_of_block(prefix, njobs);           \
  const int j = jb_.j;                                         \
  const level_job& J = jobs[j];                                \
  const int lb = jb_.lb;                                       \
  const int nb = (int)(uni(prefix[j + 1]) - uni(prefix[j]));   \
  (void)lb;                                                    \
  (void)nb

// tracker::update_input + subsample_input for every stream: the u8 pyramid of
// the new frame into the frame's slot.
__global__ __launch_bounds__(kPyrThreads) void step_pyramid_kernel(const stream_job* S, const level_job* jobs,
                                                           const frame_in* in, const uint32_t* prefix, int nstreams) {
  __shared__ pyr_args a;
  const job_block jb = job_of_block(prefix, nstreams);
  const int st = jb.j;
  const stream_job& SJ = S[st];
  const uint32_t slot = jobs[SJ.first].cnt->steps & 1u;
  const int* src = reinterpret_cast<const int*>(&SJ.pyr[slot]);
  int* dst = reinterpret_cast<int*>(&a);
  for (int i = threadIdx.x; i < (int)(sizeof(pyr_args) / 4); i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    a.frame = in[st].begin;
    a.frame_pitch = in[st].pitch;
    a.frame_vec_ok = a.frame_vec_ok && in[st].vec_ok;
  }
  __syncthreads();
  pyramid_tile<uint8_t>(a, jb.lb);
}

// Blurred-frame profile: sigma / K Gaussian of every level of the new frame
// and the Sobel/8 gradient of the blurred level (the matcher reads it as the
// previous frame's gradient next step), in one pass.
template <int K>
__global__ __launch_bounds__(64) void step_blurgrad_kernel(const level_job* jobs, const uint32_t* prefix, int njobs,
                                                           coef_pack cp) {
  CUIMG_JOB(prefix, njobs);
  const uint32_t slot = J.cnt0->steps & 1u;
  const gauss_job G = uni(J.blur[slot]);
  const grad_job g = uni(J.grad[slot]);
  gauss_sobel_tile<K>(G, g.out, g.out_pitch, lb % G.tiles_x, lb / G.tiles_x, cp);
}

// The merge scratch entries the last step's merge touched back to their idle
// state (every step, so a step after clear_scale, which skips the matcher,
// still merges on a clean scratch), then before_matching (sparse <- -1,
// particle_container.hpp:187-193).
__global__ __launch_bounds__(256) void step_fill_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  merge_restore_body(uni(J.mg[0]), J.cnt->merge_restore, lb, nb);  // J.mg[0] and [1] share the scratch
  const step_flags f = flags_of(J);
  if (!f.match) return;
  const int b = J.sparse_border;
  fill_u32_body(J.sparse, J.sparse_pitch, -b, J.nr + b, -b, J.nc + b, kEmpty, lb, nb);
}

// FAST: 7x7 windows on FP32 frames with every pixel used, which come with
// their gradient images (wave_match16).
template <typename TI, int NPL, bool SEL, bool FAST>
__global__ __launch_bounds__(kLkThreads) __attribute__((amdgpu_waves_per_eu(4))) void step_lk_kernel(const level_job* jobs, const uint32_t* prefix, int njobs,
                                                      int ppw) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.match) return;
  lk_body<TI, NPL, SEL, FAST>(J.lk[f.slot * 2 + f.cur], lb, nb, ppw);
}

__global__ __launch_bounds__(256) void step_pmask_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.detect) return;
  // the detector mask applied to the saliency: the score is per pixel, so
  // scoring every pixel and then zeroing the 3x3 block of every particle gives
  // the masked saliency
  const detect_args d = uni(J.det[f.cur]);
  particle_mask_core(d.parts, d.cnt->n, J.sal, J.sal_pitch, d.nr, d.nc, lb, nb);
}

// The bc2s strategy's detector mask also holds 0 where the contrast is below
// 5 (flow_mask_kernel, tracking_strategies.hpp:228-246): the contrast image is
// the FAST mask with threshold 5.
__global__ __launch_bounds__(256) void step_fast_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.detect) return;
  const bool bc = J.prm.bc2s;
  // the tracker's images: clamp-filled border 3, 128 bytes before column 0,
  // rows 128-byte aligned; fast_th >= 0 (checked on the host)
  fast_tile4_body(J.u8[f.slot], J.u8_pitch, bc ? J.bc.con : nullptr, bc ? J.bc.con_pitch : 0, J.sal, J.sal_pitch,
                  J.nr, J.nc, lb, J.fast_tiles_x, J.prm.fast_n, (int)J.prm.fast_th, bc ? 5 : 1);
}

// bc2s strategy: contrast and feature images of the new frame.
__global__ __launch_bounds__(256) void step_bprep_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  bc2s_prep_tile(J.bc.prep[J.cnt0->steps & 1u], lb);
}

// bc2s strategy: match_particles_kernel over the current vector.
__global__ __launch_bounds__(256) void step_bmatch_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.match) return;
  __shared__ uint32_t blk_live, blk_acc;
  if (threadIdx.x == 0) {
    blk_live = 0;
    blk_acc = 0;
  }
  __syncthreads();
  const detect_args& d = J.det[f.cur];
  const uint32_t n = J.cnt->n;
  if (lb == 0 && threadIdx.x == 0) J.cnt->n_matched = n;
  // 8 lanes per particle
  const uint32_t first = ((uint32_t)lb * blockDim.x + threadIdx.x) >> 3, stride = ((uint32_t)nb * blockDim.x) >> 3;
  uint32_t live = 0, acc = 0;
  bc2s_match_body(J.bc, d.parts, J.bc.feat[f.cur], n, J.sparse, J.sparse_pitch, J.bc.dec_accept, J.bc.dec_tr,
                  J.bc.dec_dist, first, stride, &live, &acc);
  if (live) atomicAdd(&blk_live, live);
  if (acc) atomicAdd(&blk_acc, acc);
  __syncthreads();
  if (threadIdx.x == 0 && (blk_live | blk_acc)) {
    atomicAdd(&J.cnt->live_total, (unsigned long long)blk_live);
    atomicAdd(&J.cnt->accepted_total, (unsigned long long)blk_acc);
  }
}

// bc2s strategy: per-cell flow statistics of the matched particles.
__global__ __launch_bounds__(256) void step_fstat_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  bc2s_stats_strip(J.bc, J.det[f.cur].parts, J.sparse, J.sparse_pitch, lb);
}

// bc2s strategy: the statistics fused over the scales into flow and
// multiscale count (the stream's scales are consecutive jobs).
__global__ __launch_bounds__(256) void step_fuse_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const int j0 = j - J.s;
  bc2s_fuse_body(J.bc, J.s, [&](int t) -> const bc2s_job& { return jobs[j0 + t].bc; },
                 (uint32_t)lb * blockDim.x + threadIdx.x, (uint32_t)nb * blockDim.x);
}

// bc2s strategy: filter_bad_particles_kernel over the vector after the detector.
__global__ __launch_bounds__(256) void step_filter_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.filter) return;
  const uint32_t c = f.cur ^ (f.detect ? 1u : 0u);
  bc2s_filter_body(J.bc, J.det[c].parts, J.cnt->n, (uint32_t)lb * blockDim.x + threadIdx.x,
                   (uint32_t)nb * blockDim.x);
}

// Box selection with sub-row counts, and the compaction counts.
__global__ __launch_bounds__(256) void step_detc_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.detect) return;
  const detect_args d = uni(J.det[f.cur]);
  if (lb == 0 && threadIdx.x == 0) J.cnt->n_old = J.cnt->n;  // the size before this compaction
  if (lb < d.nbr) nms_row_body(d, J.prm.fast_th, lb);
  else compact_count_body(d, lb - d.nbr);
}

// Compaction scatter and insertion of the new points (disjoint outputs).
__global__ __launch_bounds__(256) void step_scati_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.detect) return;
  const detect_args d = uni(J.det[f.cur]);
  if (lb < d.nblocks) compact_scatter_step(d, lb);
  else insert_new_step(d, lb - d.nblocks);
}

__global__ __launch_bounds__(256) void step_mscan_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.merge) return;
  const merge_args m = uni(J.mg[f.cur ^ (f.detect ? 1u : 0u)]);
  merge_scan_body(m, lb, nb);
}

__global__ __launch_bounds__(256) void step_mr1_kernel(const level_job* jobs, const uint32_t* prefix, int njobs) {
  CUIMG_JOB(prefix, njobs);
  const step_flags f = flags_of(J);
  if (!f.merge) return;
  const merge_args m = uni(J.mg[f.cur ^ (f.detect ? 1u : 0u)]);
  merge_round1_body(m, lb, nb);
}

// Merge resolution, then the end of the job's frame step.
__global__ __launch_bounds__(1024) void step_mres_kernel(const level_job* jobs) {
  const level_job& J = jobs[blockIdx.x];
  const step_flags f = flags_of(J);
  const uint32_t c = f.cur ^ (f.detect ? 1u : 0u);
  if (f.merge) {
    const merge_args m = uni(J.mg[c]);
    merge_resolve_body(m);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (!f.merge) J.cnt->merge_restore = 0;
    J.cnt->cur = c;
    J.cnt->compact_has_run = f.detect ? 1u : 0u;
    J.cnt->frame = f.frame + 1;
    if (J.s == 0) J.cnt->steps = J.cnt->steps + 1;  // read by the next step's kernels only
  }
}

// Channel 0 of a packed interleaved u8 frame (the reference sample's get_x).
__global__ __launch_bounds__(256) void extract_x_kernel(const uint8_t* raw, int channels, uint8_t* out, int pitch,
                                                        int nr, int nc) {
  const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
  if (c < nc) out[(size_t)r * pitch + c] = raw[((size_t)r * nc + c) * channels];
}

// ------------------------------------------------------------------ host
static void ingest_free(cuimg_hip_ctx* ctx, host_ingest& H) {
  for (int i = 0; i < 2; i++) {
    if (H.copied[i]) (void)hipEventSynchronize(H.copied[i]);
    if (H.consumed[i]) (void)hipEventSynchronize(H.consumed[i]);
    if (H.pinned[i]) (void)hipHostFree(H.pinned[i]);
    if (H.raw[i]) (void)hipFree(H.raw[i]);
    image_free(ctx, &H.gray[i]);
    if (H.copied[i]) (void)hipEventDestroy(H.copied[i]);
    if (H.consumed[i]) (void)hipEventDestroy(H.consumed[i]);
  }
  H = host_ingest{};
}

static int ingest_alloc(cuimg_hip_ctx* ctx, host_ingest& H, int nr, int nc, int channels) {
  ingest_free(ctx, H);
  const size_t bytes = (size_t)nr * nc * channels;
  for (int i = 0; i < 2; i++) {
    if (hipHostMalloc(reinterpret_cast<void**>(&H.pinned[i]), bytes) != hipSuccess ||
        (channels > 1 && hipMalloc(&H.raw[i], bytes) != hipSuccess) ||
        hipEventCreateWithFlags(&H.copied[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&H.consumed[i], hipEventDisableTiming) != hipSuccess) {
      ingest_free(ctx, H);
      return set_error(ctx, CUIMG_E_NOMEM, "run_host: staging allocation failed");
    }
    const int rc = image_alloc(ctx, nr, nc, 0, CUIMG_U8, &H.gray[i]);
    if (rc) {
      ingest_free(ctx, H);
      return rc;
    }
  }
  H.channels = channels;
  return 0;
}

// Stages a host frame of stream i: copy into the next pinned buffer (the call
// returns after this copy), upload on the copy stream once the step that last
// read the device frame is done, make the context stream wait for the upload,
// extract channel 0 of an interleaved frame.  *dev is the device frame.
static int stage_host_frame(cuimg_hip_batch* b, int i, const uint8_t* frame, size_t pitch, int channels,
                            const cuimg_img** dev) {
  cuimg_hip_ctx* ctx = b->ctx;
  const level_res& R0 = b->lv[b->first[i]];
  const int nr = R0.nr, nc = R0.nc;
  const size_t row = (size_t)nc * channels;
  if (!frame || (channels != 1 && channels != 3 && channels != 4))
    return set_error(ctx, CUIMG_E_INVALID, "host frame %d: null or %d channels", i, channels);
  if (pitch < row) return set_error(ctx, CUIMG_E_INVALID, "host frame %d: pitch %zu below %zu", i, pitch, row);
  host_ingest& H = b->ingest[i];
  if (H.channels != channels) CUIMG_TRY(ingest_alloc(ctx, H, nr, nc, channels));
  const int k = H.k;
  CUIMG_HIP(ctx, hipEventSynchronize(H.copied[k]));  // the upload from this staging buffer two steps ago
  if (pitch == row) {
    std::memcpy(H.pinned[k], frame, row * nr);
  } else {
    for (int r = 0; r < nr; r++) std::memcpy(H.pinned[k] + row * r, frame + pitch * r, row);
  }
  CUIMG_HIP(ctx, hipStreamWaitEvent(b->copy, H.consumed[k], 0));
  if (channels == 1)
    CUIMG_HIP(ctx, hipMemcpy2DAsync(H.gray[k].begin, H.gray[k].pitch, H.pinned[k], row, row, nr, hipMemcpyHostToDevice,
                                    b->copy));
  else
    CUIMG_HIP(ctx, hipMemcpyAsync(H.raw[k], H.pinned[k], row * nr, hipMemcpyHostToDevice, b->copy));
  CUIMG_HIP(ctx, hipEventRecord(H.copied[k], b->copy));
  CUIMG_HIP(ctx, hipStreamWaitEvent(ctx->stream, H.copied[k], 0));
  if (channels > 1)
    CUIMG_TRY(launch(ctx, "extract_x", [&] {
      extract_x_kernel<<<dim3(ceil_div(nc, 256), nr), 256, 0, ctx->stream>>>(
          static_cast<const uint8_t*>(H.raw[k]), channels, static_cast<uint8_t*>(H.gray[k].begin), H.gray[k].pitch,
          nr, nc);
    }));
  *dev = &H.gray[k];
  return 0;
}

// After the step that reads stream i's staged frame is enqueued.
static int host_frame_consumed(cuimg_hip_batch* b, int i) {
  host_ingest& H = b->ingest[i];
  CUIMG_HIP(b->ctx, hipEventRecord(H.consumed[H.k], b->ctx->stream));
  H.k ^= 1;
  return 0;
}

static int dev_alloc(cuimg_hip_batch* b, size_t bytes, void** out) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess)
    return set_error(b->ctx, CUIMG_E_NOMEM, "tracker: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b->allocs.push_back(p);
  *out = p;
  return 0;
}

template <typename T>
static int dev_alloc_n(cuimg_hip_batch* b, size_t n, T** out) {
  void* p = nullptr;
  CUIMG_TRY(dev_alloc(b, n * sizeof(T), &p));
  *out = static_cast<T*>(p);
  return 0;
}

static void destroy_batch(cuimg_hip_batch* b) {
  if (!b) return;
  drop_child(b->ctx, b);
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->exec_step) (void)hipGraphExecDestroy(b->exec_step);
  for (auto& R : b->lv) {
    for (int k = 0; k < 2; k++) {
      image_free(b->ctx, &R.u8[k]);
      image_free(b->ctx, &R.blur[k]);
      image_free(b->ctx, &R.grad[k]);
    }
    image_free(b->ctx, &R.sal);
    image_free(b->ctx, &R.sparse);
  }
  for (auto& H : b->ingest) ingest_free(b->ctx, H);
  if (b->copy) (void)hipStreamDestroy(b->copy);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->hin) (void)hipHostFree(b->hin);
  for (auto& e : b->ring_ev)
    if (e) (void)hipEventDestroy(e);
  if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
  if (b->join_ev) (void)hipEventDestroy(b->join_ev);
  if (b->aux) (void)hipStreamDestroy(b->aux);
  for (auto* v : b->views) delete v;
  delete b;
}

static int fill_sparse(cuimg_hip_ctx* ctx, level_res& R) {
  const int bd = R.sparse.border;
  return launch_fill_u32(ctx, static_cast<uint32_t*>(R.sparse.begin), R.sparse.pitch, -bd, R.nr + bd, -bd,
                         R.nc + bd, 0xFFFFFFFFu);
}

static int check_params(cuimg_hip_ctx* ctx, int nrows, int ncols, const cuimg_tracker_params* P) {
  // the matcher addresses images with 32-bit byte offsets (24-bit rows and pitch)
  if (nrows < 1 || ncols < 1 || nrows > 16384 || ncols > 16384)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: frame %dx%d (each side 1..16384)", nrows, ncols);
  if (P->nscales < 1 || P->nscales > 7)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: nscales %d not in 1..7", P->nscales);
  if (P->lk_ws < 1 || P->lk_ws > 15 || !(P->lk_ws & 1))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: LK window %d must be odd and <= 15", P->lk_ws);
  if (P->box_size < 1 || P->box_size > kMaxBox)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: detector box %d not in 1..%d", P->box_size, kMaxBox);
  if (P->detector_frequency < 1 || P->filtering_frequency < 1 || P->fast_th < 0.f || P->fast_n < 1)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: bad detector / filtering settings");
  if (P->lk_top_n < 1 || P->lk_max_iters < 0)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: LK top_n %d (>= 1) / max_iters %d (>= 0)", P->lk_top_n,
                     P->lk_max_iters);
  if (P->profile != 0 && P->profile != 1 && P->profile != 2)
    return set_error(ctx, CUIMG_E_INVALID, "tracker: profile %d not in {0,1,2}", P->profile);
  if (P->profile == 1 && (P->blur_K < 1 || P->blur_K > kMaxK || P->blur_sigma < 1))
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: blur sigma %d / half width %d", P->blur_sigma, P->blur_K);
  int r = nrows, c = ncols;
  for (int s = 0; s < P->nscales; s++) {
    if (r < 1 || c < 1)
      return set_error(ctx, CUIMG_E_INVALID, "tracker: scale %d of %dx%d is empty", s, nrows, ncols);
    if ((size_t)r * c / 4 > kMergeMaxCap)
      return set_error(ctx, CUIMG_E_UNSUPPORTED, "tracker: %dx%d exceeds the particle capacity limit", r, c);
    r /= 2;
    c /= 2;
  }
  return 0;
}

// Detector boxes of a scale: (int)(n / box) per axis (fast_detector.hpp:225).
static void set_boxes(level_res& R) {
  R.nbr = (int)((float)R.nr / (float)R.sp.box_size);
  R.nbc = (int)((float)R.nc / (float)R.sp.box_size);
}

static int alloc_level(cuimg_hip_batch* b, const cuimg_tracker_params* P, int r, int c, level_res& R) {
  cuimg_hip_ctx* ctx = b->ctx;
  R.nr = r;
  R.nc = c;
  // particle_domain_ = domain - border(3) (pyrlk_tracker.hpp:792)
  R.sp = cuimg_scale_params{P->fast_n, P->fast_th, P->detector_frequency, P->filtering_frequency, P->k,
                            P->with_merge, P->box_size, 3, 3, r - 4, c - 4};
  for (int k = 0; k < 2; k++) {
    CUIMG_TRY(image_alloc(ctx, r, c, 3, CUIMG_U8, &R.u8[k]));
    if (P->profile == 1) {
      CUIMG_TRY(image_alloc(ctx, r, c, 0, CUIMG_F32, &R.blur[k]));
      CUIMG_TRY(image_alloc(ctx, r, c, 0, CUIMG_F32X2, &R.grad[k]));
    }
  }
  CUIMG_TRY(image_alloc(ctx, r, c, 1, CUIMG_U8, &R.sal));
  CUIMG_TRY(image_alloc(ctx, r, c, 7, CUIMG_U32, &R.sparse));
  set_boxes(R);
  // detector buffers sized for the smallest box (1), so the box size of a
  // scale can change at any frame
  const size_t nboxes = (size_t)r * c;
  R.cap = (uint32_t)std::max<size_t>((size_t)r * c / 4, 4096);
  R.nblocks = ceil_div((int)R.cap, kCompactBlock);
  CUIMG_TRY(dev_alloc_n(b, std::max<size_t>(nboxes, 1), &R.boxcode));
  CUIMG_TRY(dev_alloc_n(b, std::max<size_t>((size_t)r, 1), &R.subcnt));
  for (int k = 0; k < 2; k++) CUIMG_TRY(dev_alloc_n(b, R.cap, &R.parts[k]));
  CUIMG_TRY(dev_alloc_n(b, R.cap + nboxes, &R.matches));
  CUIMG_TRY(dev_alloc_n(b, R.cap, &R.flags));
  CUIMG_TRY(dev_alloc_n(b, R.nblocks, &R.blockcnt));
  CUIMG_TRY(dev_alloc_n(b, R.cap, &R.dec_accept));
  CUIMG_TRY(dev_alloc_n(b, R.cap, &R.dec_tr));
  CUIMG_TRY(dev_alloc_n(b, R.cap, &R.dec_dist));
  if (P->profile == 2) {
    // generic_strategy: contrast_ (saturated), feature_ (s1 border 3, s2
    // border 6), descriptors, and the flow cells (domain_div_up(d, 8))
    CUIMG_TRY(image_alloc(ctx, r, c, 0, CUIMG_U8, &R.con));
    CUIMG_TRY(image_alloc(ctx, r, c, 3, CUIMG_U8, &R.s1));
    CUIMG_TRY(image_alloc(ctx, r, c, 6, CUIMG_U8, &R.s2));
    CUIMG_TRY(dev_alloc_n(b, (size_t)r * c, &R.desc));
    for (int k = 0; k < 2; k++) CUIMG_TRY(dev_alloc_n(b, R.cap, &R.feat[k]));
    R.cr = ceil_div(r, kCell);
    R.cc = ceil_div(c, kCell);
    const size_t ncell = (size_t)R.cr * R.cc;
    CUIMG_TRY(dev_alloc_n(b, ncell, &R.stats));
    CUIMG_TRY(dev_alloc_n(b, ncell, &R.flow));
    CUIMG_TRY(dev_alloc_n(b, ncell, &R.msc));
  }
  void* m = nullptr;
  CUIMG_TRY(dev_alloc(b, merge_scratch_bytes(R.cap), &m));
  R.merge = merge_scratch_carve(m, R.cap);
  CUIMG_TRY(launch_merge_reset(ctx, R.merge));
  // particle_container ctor: sparse <- -1 (tracking2/particle_container.hpp:23-31)
  CUIMG_TRY(fill_sparse(ctx, R));
  return 0;
}

static lk_level_view matcher_view(const cuimg_hip_batch* b, int j, int prev, int cur) {
  const level_res& R = b->lv[j];
  const cuimg_img& A = b->P0.profile == 0 ? R.u8[prev] : R.blur[prev];
  const cuimg_img& B = b->P0.profile == 0 ? R.u8[cur] : R.blur[cur];
  const cuimg_img& G = R.grad[prev];  // empty for the u8 profile
  return lk_level_view{A.begin, B.begin, A.pitch, B.pitch, R.nr, R.nc, static_cast<const float2*>(G.begin), G.pitch,
                       0};
}

static int blur_th(const cuimg_hip_batch* b) {
  if (b->blur_th > 0) return b->blur_th;
  // aim for >= ~4096 waves over the whole launch, 4..128 rows per tile (a
  // 1080p frame's four levels: 4 rows, 14 us against 20 us at 16 rows)
  long strips = 0;
  for (const auto& R : b->lv) strips += (long)R.nr * ceil_div(R.nc, 256);
  return (int)std::max<long>(4, std::min<long>(128, (strips + 4095) / 4096));
}

static void build_bc2s(cuimg_hip_batch* b, int i, int s, level_job& J);

// Job descriptors of job j from its resources and settings.
static void build_job(cuimg_hip_batch* b, int i, int s) {
  const int j = b->first[i] + s;
  const level_res& R = b->lv[j];
  const cuimg_tracker_params& P = b->P[i];
  level_job& J = b->hj[j];
  J = level_job{};
  J.stream = i;
  J.s = s;
  J.nr = R.nr;
  J.nc = R.nc;
  J.prm = job_params{R.sp.detector_frequency, R.sp.filtering_frequency, R.sp.with_merge, R.sp.fast_n, R.sp.fast_th,
                     P.profile == 2 ? 1 : 0};
  J.cnt = b->cnt + j;
  J.cnt0 = b->cnt + b->first[i];
  for (int k = 0; k < 2; k++) J.u8[k] = static_cast<uint8_t*>(R.u8[k].begin);
  J.u8_pitch = R.u8[0].pitch;
  J.sal = static_cast<uint8_t*>(R.sal.begin);
  J.sal_pitch = R.sal.pitch;
  J.sparse = static_cast<uint32_t*>(R.sparse.begin);
  J.sparse_pitch = R.sparse.pitch;
  J.sparse_border = R.sparse.border;
  J.fast_tiles_x = ceil_div(R.nc, kFTW);
  // fills: one workgroup per row (the fill bodies loop over rows)
  J.fill_sparse_blocks = std::max(1, std::min(1024, R.nr + 2 * R.sparse.border));
  for (int slot = 0; slot < 2; slot++)
    for (int cur = 0; cur < 2; cur++) {
      lk_args& a = J.lk[slot * 2 + cur];
      for (int k = 0; k < P.nscales; k++) a.lv[k] = matcher_view(b, b->first[i] + k, slot ^ 1, slot);
      a.nlevels = P.nscales;
      a.s = s;
      a.parts = R.parts[cur];
      a.cnt = J.cnt;
      a.sparse = J.sparse;
      a.sparse_pitch = J.sparse_pitch;
      a.dec_accept = R.dec_accept;
      a.dec_tr = R.dec_tr;
      a.dec_dist = R.dec_dist;
      a.pd_r1 = R.sp.pd_r1;
      a.pd_c1 = R.sp.pd_c1;
      a.pd_r2 = R.sp.pd_r2;
      a.pd_c2 = R.sp.pd_c2;
      a.P = make_lk_params(P.lk_ws, P.lk_max_iters, std::min(P.lk_top_n, P.lk_ws * P.lk_ws), P.lk_eps, R.sp.k,
                           P.profile == 0 ? 1 : 0);
    }
  for (int cur = 0; cur < 2; cur++) {
    detect_args& d = J.det[cur];
    d.nr = R.nr;
    d.nc = R.nc;
    d.box = R.sp.box_size;
    d.nbr = R.nbr;
    d.nbc = R.nbc;
    d.fast_th = R.sp.fast_th;
    d.sal = J.sal;
    d.sal_pitch = J.sal_pitch;
    d.sparse = J.sparse;
    d.sparse_pitch = J.sparse_pitch;
    d.parts = R.parts[cur];
    d.parts_out = R.parts[cur ^ 1];
    d.matches = R.matches;
    d.flags = R.flags;
    d.blockcnt = R.blockcnt;
    d.boxcode = R.boxcode;
    d.subcnt = R.subcnt;
    d.cnt = J.cnt;
    d.cap = R.cap;
    d.nblocks = R.nblocks;
    d.feat = R.feat[cur];
    d.feat_out = R.feat[cur ^ 1];
    d.desc = R.desc;
    d.desc_pitch = R.nc;
    merge_args& m = J.mg[cur];
    m = R.merge;
    m.nr = R.nr;
    m.nc = R.nc;
    m.parts = R.parts[cur];
    m.sparse = J.sparse;
    m.sparse_pitch = J.sparse_pitch;
    m.cnt = J.cnt;
    m.defer_restore = 1;  // step_fill_kernel resets the scratch
  }
  for (int slot = 0; slot < 2; slot++) {
    grad_job& gj = J.grad[slot];
    gj = grad_job{};
    if (P.profile != 1) continue;
    gj = grad_job{static_cast<const float*>(R.blur[slot].begin), static_cast<float2*>(R.grad[slot].begin),
                  R.blur[slot].pitch, R.grad[slot].pitch, R.nr, R.nc};
  }
  const int th = blur_th(b);
  for (int slot = 0; slot < 2; slot++) {
    gauss_job& g = J.blur[slot];
    g = gauss_job{};
    if (P.profile != 1) continue;
    g.in = R.u8[slot].begin;
    g.out = R.blur[slot].begin;
    g.in_pitch = R.u8[slot].pitch;
    g.out_pitch = R.blur[slot].pitch;
    g.nr = R.nr;
    g.nc = R.nc;
    g.th = th;
    g.tiles_x = ceil_div(R.nc, 256);
    g.vec_ok = (R.u8[slot].pitch % 4 == 0) && ((uintptr_t)R.u8[slot].begin % 4 == 0) &&
               (R.blur[slot].pitch % 16 == 0) && ((uintptr_t)R.blur[slot].begin % 16 == 0);
  }
  build_bc2s(b, i, s, J);
}

// The bc2s strategy's job of scale s (profile 2).
static void build_bc2s(cuimg_hip_batch* b, int i, int s, level_job& J) {
  const level_res& R = b->lv[b->first[i] + s];
  bc2s_job& B = J.bc;
  B = bc2s_job{};
  if (b->P[i].profile != 2) return;
  for (int slot = 0; slot < 2; slot++) {
    bc2s_prep_args& a = B.prep[slot];
    a.in = static_cast<const uint8_t*>(R.u8[slot].begin);
    a.in_pitch = R.u8[slot].pitch;
    a.con = static_cast<uint8_t*>(R.con.begin);
    a.con_pitch = R.con.pitch;
    a.con32 = nullptr;
    a.con32_pitch = 0;
    a.s1 = static_cast<uint8_t*>(R.s1.begin);
    a.s1_pitch = R.s1.pitch;
    a.s1_border = R.s1.border;
    a.s2 = static_cast<uint8_t*>(R.s2.begin);
    a.s2_pitch = R.s2.pitch;
    a.s2_border = R.s2.border;
    a.desc = R.desc;
    a.desc_pitch = R.nc;
    a.nr = R.nr;
    a.nc = R.nc;
    a.tiles_x = ceil_div(R.nc, kB2W);
  }
  B.F = bc2s_view{static_cast<const uint8_t*>(R.s1.begin), static_cast<const uint8_t*>(R.s2.begin), R.s1.pitch,
                  R.s2.pitch, R.nr, R.nc};
  B.desc = R.desc;
  B.desc_pitch = R.nc;
  B.con = static_cast<const uint8_t*>(R.con.begin);
  B.con_pitch = R.con.pitch;
  B.feat[0] = R.feat[0];
  B.feat[1] = R.feat[1];
  B.cr = R.cr;
  B.cc = R.cc;
  B.stats = R.stats;
  B.flow = R.flow;
  B.msc = R.msc;
  B.k = R.sp.k;
  B.last = b->P[i].nscales - 1;
  B.dec_accept = R.dec_accept;
  B.dec_tr = R.dec_tr;
  B.dec_dist = R.dec_dist;
}

static void build_stream(cuimg_hip_batch* b, int i) {
  stream_job& S = b->hs[i];
  S = stream_job{};
  S.first = b->first[i];
  S.nlev = b->P[i].nscales;
  for (int slot = 0; slot < 2; slot++) {
    pyr_args& a = S.pyr[slot];
    a.copy_level0 = 1;
    a.n = S.nlev;
    for (int s = 0; s < S.nlev; s++) {
      const cuimg_img& L = b->lv[S.first + s].u8[slot];
      a.lv[s] = pyr_level{L.begin, L.pitch, L.nrows, L.ncols, L.border};
    }
    const cuimg_img& L0 = b->lv[S.first].u8[slot];
    a.frame_vec_ok = 1;  // and the input table's own flag, at the step
    a.l0_vec_ok = ((uintptr_t)L0.begin % 16 == 0) && (L0.pitch % 16 == 0);
    pyr_tiling(a, L0.nrows, L0.ncols, 1);
  }
}

// Per-kind block counts of every job, their prefix sums, and the upload.
static int build_maps(cuimg_hip_batch* b) {
  std::vector<uint32_t> cnt[K_NKINDS];
  for (int i = 0; i < b->nstreams; i++) {
    const pyr_args& a = b->hs[i].pyr[0];  // both slots have the same tiling
    cnt[K_PYR].push_back((uint32_t)(a.tiles_x * ceil_div(a.lv[0].nr, 1 << a.th_log2)));
  }
  const int th = blur_th(b);
  for (int j = 0; j < b->njobs; j++) {
    const level_res& R = b->lv[j];
    const level_job& J = b->hj[j];
    const long cap = R.cap;
    cnt[K_BLUR].push_back(b->P0.profile == 1 ? (uint32_t)(ceil_div(R.nc, 256) * ceil_div(R.nr, th)) : 0u);
    cnt[K_FILL].push_back((uint32_t)J.fill_sparse_blocks);
    cnt[K_PMASK].push_back((uint32_t)std::max<long>(1, std::min<long>(512, (cap + 255) / 256)));
    cnt[K_FAST].push_back((uint32_t)(J.fast_tiles_x * ceil_div(R.nr, kFTH)));
    cnt[K_DETC].push_back((uint32_t)(R.nbr + R.nblocks));
    cnt[K_SCATI].push_back((uint32_t)(R.nblocks + ceil_div(R.nbr * R.sp.box_size, 4)));
    cnt[K_MSCAN].push_back((uint32_t)std::max<long>(1, std::min<long>(512, (cap + 255) / 256)));
    cnt[K_MR1].push_back(cnt[K_MSCAN].back());
    const bool bc = b->P0.profile == 2;
    const long cells = (long)R.cr * R.cc;
    cnt[K_BPREP].push_back(bc ? (uint32_t)(ceil_div(R.nc, kB2W) * ceil_div(R.nr, kB2H)) : 0u);
    cnt[K_BMATCH].push_back(bc ? (uint32_t)std::max<long>(1, std::min<long>(2048, (8 * cap + 255) / 256)) : 0u);
    cnt[K_FSTAT].push_back(bc ? (uint32_t)(R.cr * ceil_div(R.nc, 256)) : 0u);
    cnt[K_FUSE].push_back(bc ? (uint32_t)std::max<long>(1, (cells + 255) / 256) : 0u);
    cnt[K_FILT].push_back(bc ? (uint32_t)std::max<long>(1, std::min<long>(512, (cap + 255) / 256)) : 0u);
  }
  // Matcher: one workgroup per kLkWaves * ppw particles of capacity, at most
  // kLkMaxWgs / 4^s for scale s (the workgroups loop over the chunks, see
  // lk_body).  A coarser scale holds about a quarter of the particles of the
  // next finer one; more workgroups than that only add nearly idle ones,
  // each of which still sets up before it finds its chunks (1080p: 221 ->
  // 203 us per launch against 1024 workgroups for every scale).
  for (int j = 0; j < b->njobs; j++)
    cnt[K_LK].push_back(b->P0.profile == 2 ? 0u
                                           : (uint32_t)std::min(ceil_div((int)b->lv[j].cap, kLkWaves * b->lk_ppw),
                                                                std::max(16, kLkMaxWgs >> (2 * b->hj[j].s))));
  for (int k = 0; k < K_NKINDS; k++) {
    // prefix of block counts, then the block -> (job, block in job) map
    std::vector<uint32_t>& p = b->prefix_h[k];
    const size_t nj = cnt[k].size();
    p.assign(nj + 1, 0);
    for (size_t x = 0; x < nj; x++) p[x + 1] = p[x] + cnt[k][x];
    b->total[k] = p.back();
    if (nj >= (1u << (32 - kJobShift)))
      return set_error(b->ctx, CUIMG_E_UNSUPPORTED, "batch: %zu jobs (at most %u)", nj, 1u << (32 - kJobShift));
    for (size_t x = 0; x < nj; x++) {
      if (cnt[k][x] > kJobBlockMask + 1)
        return set_error(b->ctx, CUIMG_E_UNSUPPORTED, "batch: %u blocks in one job", cnt[k][x]);
      for (uint32_t lb = 0; lb < cnt[k][x]; lb++) p.push_back(((uint32_t)x << kJobShift) | lb);
    }
    if (p.size() > b->table_cap[k]) {
      if (b->prefix_d[k]) {
        (void)hipStreamSynchronize(b->ctx->stream);
        (void)hipFree(b->prefix_d[k]);
        b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void*)b->prefix_d[k]));
        b->prefix_d[k] = nullptr;
      }
      CUIMG_TRY(dev_alloc_n(b, p.size(), &b->prefix_d[k]));
      b->table_cap[k] = p.size();
    }
    CUIMG_HIP(b->ctx, hipMemcpyAsync(b->prefix_d[k], p.data(), p.size() * 4, hipMemcpyHostToDevice, b->ctx->stream));
  }
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

static int upload_tables(cuimg_hip_batch* b) {
  CUIMG_HIP(b->ctx, hipMemcpyAsync(b->dj, b->hj.data(), sizeof(level_job) * b->njobs, hipMemcpyHostToDevice,
                                   b->ctx->stream));
  CUIMG_HIP(b->ctx, hipMemcpyAsync(b->ds, b->hs.data(), sizeof(stream_job) * b->nstreams, hipMemcpyHostToDevice,
                                   b->ctx->stream));
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

static int rebuild(cuimg_hip_batch* b) {
  for (int i = 0; i < b->nstreams; i++) {
    for (int s = 0; s < b->P[i].nscales; s++) build_job(b, i, s);
    build_stream(b, i);
  }
  CUIMG_TRY(build_maps(b));
  CUIMG_TRY(upload_tables(b));
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->exec_step) {
    (void)hipGraphExecDestroy(b->exec_step);
    b->exec_step = nullptr;
  }
  return 0;
}

static int create_batch(cuimg_hip_ctx* ctx, int nstreams, const int* nrows, const int* ncols,
                        const cuimg_tracker_params* params, cuimg_hip_batch** out) {
  if (!ctx || !out || nstreams < 1 || !nrows || !ncols || !params) return CUIMG_E_INVALID;
  *out = nullptr;
  const cuimg_tracker_params& P0 = params[0];
  for (int i = 0; i < nstreams; i++) {
    CUIMG_TRY(check_params(ctx, nrows[i], ncols[i], &params[i]));
    const cuimg_tracker_params& P = params[i];
    if (P.profile != P0.profile || P.lk_ws != P0.lk_ws || P.lk_top_n != P0.lk_top_n ||
        (P0.profile == 1 && (P.blur_sigma != P0.blur_sigma || P.blur_K != P0.blur_K)))
      return set_error(ctx, CUIMG_E_UNSUPPORTED,
                       "batch: stream %d differs in profile / LK window / selection / blur from stream 0", i);
  }
  auto* b = new cuimg_hip_batch();
  b->ctx = ctx;
  b->nstreams = nstreams;
  b->P0 = P0;
  b->P.assign(params, params + nstreams);
  auto fail = [&](int code) {
    destroy_batch(b);
    return code;
  };
  for (int i = 0; i < nstreams; i++) {
    b->first.push_back(b->njobs);
    b->njobs += params[i].nscales;
  }
  b->lv.resize(b->njobs);
  b->hj.resize(b->njobs);
  b->hs.resize(nstreams);
  b->ingest.resize(nstreams);
  int rc = 0;
  if ((rc = dev_alloc_n(b, b->njobs, &b->cnt))) return fail(rc);
  if (hipMemsetAsync(b->cnt, 0, sizeof(level_counters) * b->njobs, ctx->stream) != hipSuccess)
    return fail(set_error(ctx, CUIMG_E_NOMEM, "batch: counter reset failed"));
  for (int i = 0; i < nstreams; i++) {
    int r = nrows[i], c = ncols[i];
    for (int s = 0; s < params[i].nscales; s++) {
      if ((rc = alloc_level(b, &params[i], r, c, b->lv[b->first[i] + s]))) return fail(rc);
      r /= 2;
      c /= 2;
    }
  }
  if (P0.profile == 1) gaussian_coefs(P0.blur_sigma, P0.blur_K, b->cp.c);
  if ((rc = dev_alloc_n(b, b->njobs, &b->dj))) return fail(rc);
  if ((rc = dev_alloc_n(b, nstreams, &b->ds))) return fail(rc);
  if ((rc = dev_alloc_n(b, nstreams, &b->din))) return fail(rc);
  if (hipHostMalloc(reinterpret_cast<void**>(&b->hin), sizeof(frame_in) * nstreams * kRing) != hipSuccess)
    return fail(set_error(ctx, CUIMG_E_NOMEM, "batch: pinned input table allocation failed"));
  for (auto& e : b->ring_ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(ctx, CUIMG_E_NOMEM, "batch: event creation failed"));
  if (hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->copy, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(ctx, CUIMG_E_NOMEM, "batch: stream or event creation failed"));
  if ((rc = rebuild(b))) return fail(rc);
  *out = b;
  return 0;
}

static void launch_blur(cuimg_hip_batch* b) {
  const dim3 grid(b->total[K_BLUR]);
  hipStream_t st = b->ctx->stream;
  switch (b->P0.blur_K) {
#define CUIMG_K(k) \
  case k: step_blurgrad_kernel<k><<<grid, 64, 0, st>>>(b->dj, b->prefix_d[K_BLUR], b->njobs, b->cp); break;
    CUIMG_K(1) CUIMG_K(2) CUIMG_K(3) CUIMG_K(4) CUIMG_K(5) CUIMG_K(6) CUIMG_K(7) CUIMG_K(8)
    CUIMG_K(9) CUIMG_K(10) CUIMG_K(11) CUIMG_K(12) CUIMG_K(13) CUIMG_K(14) CUIMG_K(15) CUIMG_K(16)
#undef CUIMG_K
  }
}

// The fixed launch sequence of one frame step of every stream.
static int enqueue_pyramid(cuimg_hip_batch* b) {
  hipStream_t st = b->ctx->stream;
  return launch(b->ctx, "pyramid_build", [&] {
    step_pyramid_kernel<<<b->total[K_PYR], kPyrThreads, 0, st>>>(b->ds, b->dj, b->din, b->prefix_d[K_PYR], b->nstreams);
  });
}

// with_pyramid: the step builds the pyramid of the frames in the input table
// first (tracker::run(in)); otherwise it steps on the pyramid already in the
// frame slot (tracker::run()).
// generic_strategy::update of every scale (bc2s descriptor): feature images,
// FAST beside the matcher, flow statistics and fusion, detector, filter, merge.
static int enqueue_step_bc2s(cuimg_hip_batch* b, bool capturing) {
  cuimg_hip_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const int nj = b->njobs;
  const level_job* J = b->dj;
  uint32_t* const* pf = b->prefix_d;
  CUIMG_TRY(launch(ctx, "bc2s_images", [&] { step_bprep_kernel<<<b->total[K_BPREP], 256, 0, st>>>(J, pf[K_BPREP], nj); }));
  const bool fork = capturing;
  hipStream_t fst = fork ? b->aux : st;
  if (fork) {
    CUIMG_HIP(ctx, hipEventRecord(b->fork_ev, st));
    CUIMG_HIP(ctx, hipStreamWaitEvent(b->aux, b->fork_ev, 0));
  }
  CUIMG_TRY(launch(ctx, "fast_saliency", [&] { step_fast_kernel<<<b->total[K_FAST], 256, 0, fst>>>(J, pf[K_FAST], nj); }));
  if (fork) CUIMG_HIP(ctx, hipEventRecord(b->join_ev, b->aux));
  CUIMG_TRY(launch(ctx, "fills", [&] { step_fill_kernel<<<b->total[K_FILL], 256, 0, st>>>(J, pf[K_FILL], nj); }));
  CUIMG_TRY(launch(ctx, "bc2s_match", [&] { step_bmatch_kernel<<<b->total[K_BMATCH], 256, 0, st>>>(J, pf[K_BMATCH], nj); }));
  CUIMG_TRY(launch(ctx, "flow_stats", [&] { step_fstat_kernel<<<b->total[K_FSTAT], 256, 0, st>>>(J, pf[K_FSTAT], nj); }));
  CUIMG_TRY(launch(ctx, "flow_fusion", [&] { step_fuse_kernel<<<b->total[K_FUSE], 256, 0, st>>>(J, pf[K_FUSE], nj); }));
  if (fork) CUIMG_HIP(ctx, hipStreamWaitEvent(st, b->join_ev, 0));
  CUIMG_TRY(launch(ctx, "particle_mask",
                   [&] { step_pmask_kernel<<<b->total[K_PMASK], 256, 0, st>>>(J, pf[K_PMASK], nj); }));
  CUIMG_TRY(launch(ctx, "detect_count", [&] { step_detc_kernel<<<b->total[K_DETC], 256, 0, st>>>(J, pf[K_DETC], nj); }));
  CUIMG_TRY(launch(ctx, "scatter_insert",
                   [&] { step_scati_kernel<<<b->total[K_SCATI], 256, 0, st>>>(J, pf[K_SCATI], nj); }));
  CUIMG_TRY(launch(ctx, "filter", [&] { step_filter_kernel<<<b->total[K_FILT], 256, 0, st>>>(J, pf[K_FILT], nj); }));
  CUIMG_TRY(launch(ctx, "merge_scan", [&] { step_mscan_kernel<<<b->total[K_MSCAN], 256, 0, st>>>(J, pf[K_MSCAN], nj); }));
  CUIMG_TRY(launch(ctx, "merge_round1", [&] { step_mr1_kernel<<<b->total[K_MR1], 256, 0, st>>>(J, pf[K_MR1], nj); }));
  return launch(ctx, "merge_resolve", [&] { step_mres_kernel<<<nj, 1024, 0, st>>>(J); });
}

static int enqueue_step(cuimg_hip_batch* b, bool capturing, bool with_pyramid = true) {
  cuimg_hip_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const int nj = b->njobs;
  const level_job* J = b->dj;
  uint32_t* const* pf = b->prefix_d;
  if (with_pyramid) CUIMG_TRY(enqueue_pyramid(b));
  if (b->P0.profile == 2) return enqueue_step_bc2s(b, capturing);
  // FAST needs only the pyramid: in a captured step it runs beside the matcher
  // on the second stream; plain launches (and per-launch timing on the context
  // stream) keep it in order, where two more stream calls per step would only
  // add host time
  const bool fork = capturing;
  hipStream_t fst = fork ? b->aux : st;
  if (fork) {
    CUIMG_HIP(ctx, hipEventRecord(b->fork_ev, st));
    CUIMG_HIP(ctx, hipStreamWaitEvent(b->aux, b->fork_ev, 0));
  }
  CUIMG_TRY(launch(ctx, "fast_saliency", [&] { step_fast_kernel<<<b->total[K_FAST], 256, 0, fst>>>(J, pf[K_FAST], nj); }));
  // the sparse reset only has to precede the matcher: it runs on FAST's branch
  CUIMG_TRY(launch(ctx, "fills", [&] { step_fill_kernel<<<b->total[K_FILL], 256, 0, fst>>>(J, pf[K_FILL], nj); }));
  if (fork) CUIMG_HIP(ctx, hipEventRecord(b->join_ev, b->aux));
  if (b->total[K_BLUR]) CUIMG_TRY(launch(ctx, "blur_gradient", [&] { launch_blur(b); }));
  // join before the matcher (the fill's sparse image is the matcher's); a FAST
  // still running would hold CU slots the matcher's workgroups wait for
  if (fork) CUIMG_HIP(ctx, hipStreamWaitEvent(st, b->join_ev, 0));
  const bool fast = b->P0.profile == 1 && b->P0.lk_ws == 7 && b->P0.lk_top_n >= 49;
  CUIMG_TRY(launch(ctx, "lk_track", [&] {
    if (fast) {
      step_lk_kernel<float, 1, false, true><<<b->total[K_LK], kLkThreads, 0, st>>>(J, pf[K_LK], nj, b->lk_ppw);
      return;
    }
    lk_variant(b->P0.lk_ws, b->P0.lk_top_n, b->P0.profile == 0, [&](auto ti, auto npl, auto sel) {
      using TI = typename decltype(ti)::type;
      step_lk_kernel<TI, decltype(npl)::value, decltype(sel)::value, false>
          <<<b->total[K_LK], kLkThreads, 0, st>>>(J, pf[K_LK], nj, b->lk_ppw);
    });
  }));
  CUIMG_TRY(launch(ctx, "particle_mask",
                   [&] { step_pmask_kernel<<<b->total[K_PMASK], 256, 0, st>>>(J, pf[K_PMASK], nj); }));
  CUIMG_TRY(launch(ctx, "detect_count", [&] { step_detc_kernel<<<b->total[K_DETC], 256, 0, st>>>(J, pf[K_DETC], nj); }));
  CUIMG_TRY(launch(ctx, "scatter_insert",
                   [&] { step_scati_kernel<<<b->total[K_SCATI], 256, 0, st>>>(J, pf[K_SCATI], nj); }));
  CUIMG_TRY(launch(ctx, "merge_scan", [&] { step_mscan_kernel<<<b->total[K_MSCAN], 256, 0, st>>>(J, pf[K_MSCAN], nj); }));
  CUIMG_TRY(launch(ctx, "merge_round1", [&] { step_mr1_kernel<<<b->total[K_MR1], 256, 0, st>>>(J, pf[K_MR1], nj); }));
  return launch(ctx, "merge_resolve", [&] { step_mres_kernel<<<nj, 1024, 0, st>>>(J); });
}

static int read_counters(cuimg_hip_batch* b, int j, level_counters* c) {
  CUIMG_HIP(b->ctx, hipMemcpyAsync(c, b->cnt + j, sizeof(level_counters), hipMemcpyDeviceToHost, b->ctx->stream));
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

// Writes the frames of the next step into the device input table (through
// the pinned ring).
static int write_inputs(cuimg_hip_batch* b, const cuimg_img* frames) {
  cuimg_hip_ctx* ctx = b->ctx;
  const int k = (int)(b->steps % kRing);
  frame_in* row = b->hin + (size_t)k * b->nstreams;
  CUIMG_HIP(ctx, hipEventSynchronize(b->ring_ev[k]));  // the table copied kRing steps ago has been read
  for (int i = 0; i < b->nstreams; i++) {
    const cuimg_img& f = frames[i];
    const level_res& R0 = b->lv[b->first[i]];
    if (f.dtype != CUIMG_U8 || f.nrows != R0.nr || f.ncols != R0.nc || !f.begin)
      return set_error(ctx, CUIMG_E_INVALID, "step: frame %d must be U8 %dx%d", i, R0.nr, R0.nc);
    row[i] = frame_in{f.begin, f.pitch, ((uintptr_t)f.begin % 16 == 0) && (f.pitch % 16 == 0)};
  }
  CUIMG_HIP(ctx, hipMemcpyAsync(b->din, row, sizeof(frame_in) * b->nstreams, hipMemcpyHostToDevice, ctx->stream));
  CUIMG_HIP(ctx, hipEventRecord(b->ring_ev[k], ctx->stream));
  return 0;
}

// One frame step, replayed from its captured graph (or launched).
static int launch_step(cuimg_hip_batch* b, bool with_pyramid) {
  cuimg_hip_ctx* ctx = b->ctx;
  if (b->use_graph && !ctx->profiling) {
    hipGraphExec_t& exec = with_pyramid ? b->exec : b->exec_step;
    if (!exec) {
      hipGraph_t g = nullptr;
      CUIMG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_step(b, true, with_pyramid);
      const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      CUIMG_HIP(ctx, e);
      const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      CUIMG_HIP(ctx, ei);
    }
    CUIMG_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
  } else {
    CUIMG_TRY(enqueue_step(b, false, with_pyramid));
  }
  b->steps++;
  b->input_ready = false;
  return 0;
}

static int batch_step(cuimg_hip_batch* b, const cuimg_img* frames) {
  if (!b || !frames) return CUIMG_E_INVALID;
  CUIMG_TRY(write_inputs(b, frames));
  return launch_step(b, true);
}

// tracker::update_input: the pyramid of the frames into the next frame slot.
static int batch_update_input(cuimg_hip_batch* b, const cuimg_img* frames) {
  if (!b || !frames) return CUIMG_E_INVALID;
  CUIMG_TRY(write_inputs(b, frames));
  CUIMG_TRY(enqueue_pyramid(b));
  b->input_ready = true;
  return 0;
}

// tracker::run(): without a newly staged frame the last input is used again
// (its pyramid is copied into the next frame slot, borders included).
static int batch_run_step(cuimg_hip_batch* b) {
  if (!b) return CUIMG_E_INVALID;
  if (!b->input_ready) {
    for (int i = 0; i < b->nstreams; i++) {
      level_counters c;
      CUIMG_TRY(read_counters(b, b->first[i], &c));
      const int slot = (int)(c.steps & 1u);
      for (int s = 0; s < b->P[i].nscales; s++) {
        const level_res& R = b->lv[b->first[i] + s];
        const cuimg_img& from = R.u8[slot ^ 1];
        const cuimg_img& to = R.u8[slot];
        const size_t off = (size_t)from.border * from.pitch + from.border;
        CUIMG_HIP(b->ctx, hipMemcpyAsync(static_cast<uint8_t*>(to.begin) - off, static_cast<const uint8_t*>(from.begin) - off,
                                         (size_t)from.pitch * (from.nrows + 2 * from.border), hipMemcpyDeviceToDevice,
                                         b->ctx->stream));
      }
    }
  }
  return launch_step(b, false);
}


static int job_of(cuimg_hip_tracker* t, int s, int* j) {
  if (!t || !t->b) return CUIMG_E_INVALID;
  const int L = t->b->P[t->stream].nscales;
  if (s < 0 || s >= L) return set_error(t->b->ctx, CUIMG_E_INVALID, "tracker: scale %d not in 0..%d", s, L - 1);
  *j = t->b->first[t->stream] + s;
  return 0;
}

static int clear_stream(cuimg_hip_batch* b, int i) {
  for (int s = 0; s < b->P[i].nscales; s++) {
    level_res& R = b->lv[b->first[i] + s];
    CUIMG_TRY(fill_sparse(b->ctx, R));
    CUIMG_TRY(launch_merge_reset(b->ctx, R.merge));
  }
  CUIMG_HIP(b->ctx, hipMemsetAsync(b->cnt + b->first[i], 0, sizeof(level_counters) * b->P[i].nscales, b->ctx->stream));
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

// ------------------------------------------- standalone detector entry points
// Descriptors of n particles at their (truncated) positions (bc2s set_particles).
__global__ __launch_bounds__(256) void features_at_kernel(const cuimg_particle_f* parts, uint32_t n, const uint4* D,
                                                          int nr, int nc, uint4* feat) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int r = (int)parts[i].pos_r, c = (int)parts[i].pos_c;
  feat[i] = (r >= 0 && r < nr && c >= 0 && c < nc) ? D[(size_t)r * nc + c] : make_uint4(0, 0, 0, 0);
}

__global__ __launch_bounds__(256) void mask_fill_kernel(uint8_t* m, int pitch, int nr, int nc) {
  fill_u8_body(m, pitch, nr, nc, 0xFFFFFFFFu, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void mask_particles_kernel(const cuimg_particle_f* parts, uint32_t n, uint8_t* m,
                                                             int pitch, int nr, int nc) {
  particle_mask_core(parts, n, m, pitch, nr, nc, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void select_rows_kernel(detect_args d) { nms_row_body(d, d.fast_th, blockIdx.x); }

__global__ __launch_bounds__(1024) void select_scan_kernel(detect_args d) { compact_scan_body(d); }

__global__ __launch_bounds__(256) void select_points_kernel(detect_args d, int32_t* pts, uint32_t cap) {
  select_points_body(d, pts, cap, blockIdx.x);
}

// Pixel claims with the serial last-writer rule of the container: after
// claim_zero_kernel and then claim_max_kernel over the same particles [i0, i1)
// (live ones only if live_only), each of their pixels holds the largest index
// of those particles at that pixel.  Positions outside the domain claim nothing.
__global__ __launch_bounds__(256) void claim_zero_kernel(const cuimg_particle_f* parts, uint32_t i0, uint32_t i1,
                                                         int live_only, uint32_t* sp, int pitch, int nr, int nc) {
  for (uint32_t i = i0 + blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += gridDim.x * blockDim.x) {
    const cuimg_particle_f p = parts[i];
    const int r = tzf(p.pos_r), c = tzf(p.pos_c);
    if ((p.age != 0 || !live_only) && r >= 0 && r < nr && c >= 0 && c < nc) *px<uint32_t>(sp, pitch, r, c) = 0;
  }
}

__global__ __launch_bounds__(256) void claim_max_kernel(const cuimg_particle_f* parts, uint32_t i0, uint32_t i1,
                                                        int live_only, uint32_t* sp, int pitch, int nr, int nc) {
  for (uint32_t i = i0 + blockIdx.x * blockDim.x + threadIdx.x; i < i1; i += gridDim.x * blockDim.x) {
    const cuimg_particle_f p = parts[i];
    const int r = tzf(p.pos_r), c = tzf(p.pos_c);
    if ((p.age != 0 || !live_only) && r >= 0 && r < nr && c >= 0 && c < nc) atomicMax(px<uint32_t>(sp, pitch, r, c), i);
  }
}

// particle_container::add: particle n + k of age 1 and zero speed at point k.
__global__ __launch_bounds__(256) void add_points_kernel(cuimg_particle_f* parts, uint32_t n, uint32_t m,
                                                         const int32_t* pts) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
    cuimg_particle_f p;
    p.speed_r = 0.f;
    p.speed_c = 0.f;
    p.pos_r = (float)pts[2 * (size_t)k];
    p.pos_c = (float)pts[2 * (size_t)k + 1];
    p.age = 1;
    p.pad = 0;
    parts[n + k] = p;
  }
}

__global__ __launch_bounds__(256) void compact_count_kernel(detect_args d) { compact_count_body(d, blockIdx.x); }
__global__ __launch_bounds__(1024) void compact_scan_kernel(detect_args d) { compact_scan_body(d); }
__global__ __launch_bounds__(256) void compact_scatter_kernel(detect_args d) { compact_scatter_body(d, blockIdx.x); }

static int launch_claims(cuimg_hip_ctx* ctx, const cuimg_particle_f* parts, uint32_t i0, uint32_t i1, int live_only,
                         const cuimg_img* sparse) {
  if (i1 <= i0) return 0;
  const int grid = std::min(ceil_div((int)std::min<uint32_t>(i1 - i0, 1u << 30), 256), 4096);
  uint32_t* sp = static_cast<uint32_t*>(sparse->begin);
  const int pitch = (int)sparse->pitch, nr = sparse->nrows, nc = sparse->ncols;
  hipStream_t st = ctx->stream;
  CUIMG_TRY(launch(ctx, "claim_zero", [&] {
    claim_zero_kernel<<<grid, 256, 0, st>>>(parts, i0, i1, live_only, sp, pitch, nr, nc);
  }));
  return launch(ctx, "claim_max", [&] {
    claim_max_kernel<<<grid, 256, 0, st>>>(parts, i0, i1, live_only, sp, pitch, nr, nc);
  });
}

}  // namespace cuimg

using namespace cuimg;

extern "C" {

// ------------------------------------------------------------------ batch
int cuimg_hip_batch_create(cuimg_hip_ctx* ctx, int nstreams, const int* nrows, const int* ncols,
                           const cuimg_tracker_params* params, cuimg_hip_batch** out) {
  CUIMG_TRY(create_batch(ctx, nstreams, nrows, ncols, params, out));
  add_child(ctx, *out, [](void* p) { destroy_batch(static_cast<cuimg_hip_batch*>(p)); });
  return 0;
}

int cuimg_hip_batch_destroy(cuimg_hip_batch* b) {
  destroy_batch(b);
  return 0;
}

int cuimg_hip_batch_step(cuimg_hip_batch* b, const cuimg_img* frames) { return batch_step(b, frames); }

int cuimg_hip_batch_step_host(cuimg_hip_batch* b, const uint8_t* const* frames, const size_t* pitches,
                              int channels) {
  if (!b || !frames || !pitches) return CUIMG_E_INVALID;
  std::vector<cuimg_img> dev(b->nstreams);
  for (int i = 0; i < b->nstreams; i++) {
    const cuimg_img* d = nullptr;
    CUIMG_TRY(stage_host_frame(b, i, frames[i], pitches[i], channels, &d));
    dev[i] = *d;
  }
  const int rc = batch_step(b, dev.data());
  for (int i = 0; i < b->nstreams; i++) CUIMG_TRY(host_frame_consumed(b, i));
  return rc;
}

int cuimg_hip_batch_set_graph(cuimg_hip_batch* b, int enable) {
  if (!b) return CUIMG_E_INVALID;
  b->use_graph = enable != 0;
  return 0;
}

int cuimg_hip_batch_tracker(cuimg_hip_batch* b, int stream, cuimg_hip_tracker** out) {
  if (!b || !out) return CUIMG_E_INVALID;
  if (stream < 0 || stream >= b->nstreams)
    return set_error(b->ctx, CUIMG_E_INVALID, "batch: stream %d not in 0..%d", stream, b->nstreams - 1);
  auto* t = new cuimg_hip_tracker();
  t->b = b;
  t->stream = stream;
  b->views.push_back(t);
  *out = t;
  return 0;
}

int cuimg_hip_batch_set_tuning(cuimg_hip_batch* b, int lk_ppw, int blur_th) {
  if (!b || lk_ppw < 1 || lk_ppw > 256 || blur_th < 0 || blur_th > 1024) return CUIMG_E_INVALID;
  b->lk_ppw = lk_ppw;
  b->blur_th = blur_th;
  return rebuild(b);
}

int cuimg_hip_batch_get_tuning(cuimg_hip_batch* b, int* lk_ppw, int* blur_th) {
  if (!b) return CUIMG_E_INVALID;
  if (lk_ppw) *lk_ppw = b->lk_ppw;
  if (blur_th) *blur_th = ::cuimg::blur_th(b);
  return 0;
}

int cuimg_hip_batch_clear(cuimg_hip_batch* b) {
  if (!b) return CUIMG_E_INVALID;
  for (int i = 0; i < b->nstreams; i++) CUIMG_TRY(clear_stream(b, i));
  return 0;
}

// ---------------------------------------------------------------- tracker
int cuimg_hip_tracker_create(cuimg_hip_ctx* ctx, int nrows, int ncols, const cuimg_tracker_params* params,
                             cuimg_hip_tracker** out) {
  if (!out) return CUIMG_E_INVALID;
  *out = nullptr;
  cuimg_hip_batch* b = nullptr;
  CUIMG_TRY(create_batch(ctx, 1, &nrows, &ncols, params, &b));
  auto* t = new cuimg_hip_tracker();
  t->b = b;
  t->stream = 0;
  t->owner = true;
  add_child(ctx, t, [](void* p) {
    auto* tr = static_cast<cuimg_hip_tracker*>(p);
    destroy_batch(tr->b);
    delete tr;
  });
  *out = t;
  return 0;
}

int cuimg_hip_tracker_destroy(cuimg_hip_tracker* t) {
  if (!t) return 0;
  if (t->owner) {
    drop_child(t->b->ctx, t);
    destroy_batch(t->b);
    delete t;
  }
  return 0;  // views belong to their batch
}

int cuimg_hip_tracker_run(cuimg_hip_tracker* t, const cuimg_img* frame) {
  if (!t || !frame) return CUIMG_E_INVALID;
  if (t->b->nstreams != 1)
    return set_error(t->b->ctx, CUIMG_E_INVALID, "tracker_run: a stream of a batch steps with batch_step");
  return batch_step(t->b, frame);
}

int cuimg_hip_tracker_run_host_rgb(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch, int channels) {
  if (!t || !frame || (channels != 1 && channels != 3 && channels != 4)) return CUIMG_E_INVALID;
  cuimg_hip_batch* b = t->b;
  if (b->nstreams != 1)
    return set_error(b->ctx, CUIMG_E_INVALID, "tracker_run_host: a stream of a batch steps with batch_step");
  const cuimg_img* dev = nullptr;
  CUIMG_TRY(stage_host_frame(b, t->stream, frame, pitch, channels, &dev));
  const int rc = cuimg_hip_tracker_run(t, dev);
  CUIMG_TRY(host_frame_consumed(b, t->stream));
  return rc;
}

int cuimg_hip_tracker_run_host(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch) {
  return cuimg_hip_tracker_run_host_rgb(t, frame, pitch, 1);
}

int cuimg_hip_tracker_update_input(cuimg_hip_tracker* t, const cuimg_img* frame) {
  if (!t || !frame) return CUIMG_E_INVALID;
  if (t->b->nstreams != 1)
    return set_error(t->b->ctx, CUIMG_E_INVALID, "tracker_update_input: a stream of a batch steps with batch_step");
  return batch_update_input(t->b, frame);
}

int cuimg_hip_tracker_update_input_host(cuimg_hip_tracker* t, const uint8_t* frame, size_t pitch) {
  if (!t || !frame) return CUIMG_E_INVALID;
  cuimg_hip_batch* b = t->b;
  if (b->nstreams != 1)
    return set_error(b->ctx, CUIMG_E_INVALID, "tracker_update_input: a stream of a batch steps with batch_step");
  const cuimg_img* dev = nullptr;
  CUIMG_TRY(stage_host_frame(b, t->stream, frame, pitch, 1, &dev));
  const int rc = batch_update_input(b, dev);
  CUIMG_TRY(host_frame_consumed(b, t->stream));
  return rc;
}

int cuimg_hip_tracker_step(cuimg_hip_tracker* t) {
  if (!t) return CUIMG_E_INVALID;
  if (t->b->nstreams != 1)
    return set_error(t->b->ctx, CUIMG_E_INVALID, "tracker_step: a stream of a batch steps with batch_step");
  return batch_run_step(t->b);
}

int cuimg_hip_tracker_clear_scale(cuimg_hip_tracker* t, int scale) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  // level_counters::frame of the scale (its particles and the stream's step
  // count, which picks the frame slots, stay)
  cuimg_hip_batch* b = t->b;
  CUIMG_HIP(b->ctx, hipMemsetAsync(&b->cnt[j].frame, 0, sizeof(uint32_t), b->ctx->stream));
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_set_tuning(cuimg_hip_tracker* t, int lk_ppw, int blur_th) {
  if (!t || !t->b) return CUIMG_E_INVALID;
  if (t->b->nstreams != 1)
    return set_error(t->b->ctx, CUIMG_E_INVALID, "tracker_set_tuning: a stream of a batch is tuned with its batch");
  return cuimg_hip_batch_set_tuning(t->b, lk_ppw, blur_th);
}

int cuimg_hip_tracker_set_graph(cuimg_hip_tracker* t, int enable) {
  if (!t) return CUIMG_E_INVALID;
  return cuimg_hip_batch_set_graph(t->b, enable);
}

int cuimg_hip_tracker_set_scale_params(cuimg_hip_tracker* t, int scale, const cuimg_scale_params* in) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  cuimg_hip_batch* b = t->b;
  if (!in || in->fast_n < 1 || in->fast_th < 0.f || in->detector_frequency < 1 || in->filtering_frequency < 1)
    return set_error(b->ctx, CUIMG_E_INVALID, "set_scale_params: bad settings");
  if (in->box_size < 1 || in->box_size > kMaxBox)
    return set_error(b->ctx, CUIMG_E_UNSUPPORTED, "set_scale_params: detector box %d not in 1..%d", in->box_size,
                     kMaxBox);
  level_res& R = b->lv[j];
  const bool new_box = in->box_size != R.sp.box_size;
  R.sp = *in;
  if (new_box) {
    // the box size changes the launch shape of the detector kernels
    set_boxes(R);
    CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
    return rebuild(b);
  }
  build_job(b, t->stream, scale);
  CUIMG_HIP(b->ctx, hipMemcpyAsync(b->dj + j, &b->hj[j], sizeof(level_job), hipMemcpyHostToDevice, b->ctx->stream));
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_get_scale_params(cuimg_hip_tracker* t, int scale, cuimg_scale_params* out) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (!out) return CUIMG_E_INVALID;
  *out = t->b->lv[j].sp;
  return 0;
}

int cuimg_hip_tracker_clear(cuimg_hip_tracker* t) {
  if (!t || !t->b) return CUIMG_E_INVALID;
  return clear_stream(t->b, t->stream);
}

int cuimg_hip_tracker_nparticles(cuimg_hip_tracker* t, int scale, uint32_t* n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  if (n) *n = c.n;
  if (c.overflow)
    return set_error(t->b->ctx, CUIMG_E_CAPACITY, "tracker: scale %d exceeded its particle capacity %u", scale,
                     t->b->lv[j].cap);
  return 0;
}

int cuimg_hip_tracker_get_particles(cuimg_hip_tracker* t, int scale, cuimg_particle_f* out, uint32_t cap,
                                    uint32_t* n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  if (n) *n = c.n;
  if (!out) return 0;
  if (c.n > cap)
    return set_error(t->b->ctx, CUIMG_E_CAPACITY, "get_particles: %u particles, buffer holds %u", c.n, cap);
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(out, t->b->lv[j].parts[c.cur & 1], sizeof(cuimg_particle_f) * c.n,
                                      hipMemcpyDeviceToHost, t->b->ctx->stream));
  CUIMG_HIP(t->b->ctx, hipStreamSynchronize(t->b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_set_particles(cuimg_hip_tracker* t, int scale, const cuimg_particle_f* in, uint32_t n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  cuimg_hip_batch* b = t->b;
  level_res& R = b->lv[j];
  if (n > R.cap) return set_error(b->ctx, CUIMG_E_CAPACITY, "set_particles: %u exceeds capacity %u", n, R.cap);
  level_counters c;
  CUIMG_TRY(read_counters(b, j, &c));
  if (n)
    CUIMG_HIP(b->ctx, hipMemcpyAsync(R.parts[c.cur & 1], in, sizeof(cuimg_particle_f) * n, hipMemcpyHostToDevice,
                                     b->ctx->stream));
  CUIMG_HIP(b->ctx, hipMemcpyAsync(&b->cnt[j].n, &n, sizeof(uint32_t), hipMemcpyHostToDevice, b->ctx->stream));
  // bc2s strategy: each particle's descriptor at its position in the scale's
  // current feature images (zero before the first frame)
  if (b->P0.profile == 2 && n) {
    level_counters c0;
    CUIMG_TRY(read_counters(b, b->first[t->stream], &c0));
    if (c0.steps == 0) {
      CUIMG_HIP(b->ctx, hipMemsetAsync(R.feat[c.cur & 1], 0, sizeof(uint4) * n, b->ctx->stream));
    } else {
      CUIMG_TRY(launch(b->ctx, "features_at", [&] {
        features_at_kernel<<<ceil_div((int)n, 256), 256, 0, b->ctx->stream>>>(R.parts[c.cur & 1], n, R.desc, R.nr,
                                                                               R.nc, R.feat[c.cur & 1]);
      }));
    }
  }
  CUIMG_HIP(b->ctx, hipStreamSynchronize(b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_get_features(cuimg_hip_tracker* t, int scale, uint8_t* out, uint32_t cap, uint32_t* n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (t->b->P0.profile != 2) return set_error(t->b->ctx, CUIMG_E_INVALID, "get_features: not a bc2s tracker");
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  if (n) *n = c.n;
  if (!out) return 0;
  if (c.n > cap)
    return set_error(t->b->ctx, CUIMG_E_CAPACITY, "get_features: %u particles, buffer holds %u", c.n, cap);
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(out, t->b->lv[j].feat[c.cur & 1], 16 * (size_t)c.n, hipMemcpyDeviceToHost,
                                      t->b->ctx->stream));
  CUIMG_HIP(t->b->ctx, hipStreamSynchronize(t->b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_bc2s_images(cuimg_hip_tracker* t, int scale, cuimg_img* s1, cuimg_img* s2,
                                  cuimg_img* contrast) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (t->b->P0.profile != 2) return set_error(t->b->ctx, CUIMG_E_INVALID, "bc2s_images: not a bc2s tracker");
  const level_res& R = t->b->lv[j];
  if (s1) *s1 = R.s1;
  if (s2) *s2 = R.s2;
  if (contrast) *contrast = R.con;
  return 0;
}

int cuimg_hip_tracker_flow(cuimg_hip_tracker* t, int scale, cuimg_img* flow, cuimg_img* msc) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (t->b->P0.profile != 2) return set_error(t->b->ctx, CUIMG_E_INVALID, "flow: not a bc2s tracker");
  const level_res& R = t->b->lv[j];
  if (flow) *flow = cuimg_img{R.flow, R.cr, R.cc, R.cc * 8, 0, CUIMG_F32X2};
  if (msc) *msc = cuimg_img{R.msc, R.cr, R.cc, R.cc * 4, 0, CUIMG_U32};
  return 0;
}

int cuimg_hip_tracker_get_matches(cuimg_hip_tracker* t, int scale, uint32_t* out, uint32_t cap, uint32_t* n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  const uint32_t m = c.n_old + c.n_new;
  if (n) *n = m;
  if (!out) return 0;
  if (m > cap) return set_error(t->b->ctx, CUIMG_E_CAPACITY, "get_matches: %u entries, buffer holds %u", m, cap);
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(out, t->b->lv[j].matches, sizeof(uint32_t) * m, hipMemcpyDeviceToHost,
                                      t->b->ctx->stream));
  CUIMG_HIP(t->b->ctx, hipStreamSynchronize(t->b->ctx->stream));
  return 0;
}

int cuimg_hip_tracker_compact_has_run(cuimg_hip_tracker* t, int scale, int* out) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (!out) return CUIMG_E_INVALID;
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  *out = c.compact_has_run ? 1 : 0;
  return 0;
}

int cuimg_hip_tracker_sparse(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (!view) return CUIMG_E_INVALID;
  *view = t->b->lv[j].sparse;
  return 0;
}

int cuimg_hip_tracker_level(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (!view) return CUIMG_E_INVALID;
  level_counters c;
  CUIMG_TRY(read_counters(t->b, t->b->first[t->stream], &c));
  *view = t->b->lv[j].u8[(c.steps + 1) & 1];  // slot of the last frame
  return 0;
}

int cuimg_hip_tracker_saliency(cuimg_hip_tracker* t, int scale, cuimg_img* view) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  if (!view) return CUIMG_E_INVALID;
  *view = t->b->lv[j].sal;
  return 0;
}

int cuimg_hip_tracker_get_decisions(cuimg_hip_tracker* t, int scale, int32_t* accept, float* tr2, float* dist,
                                    uint32_t cap, uint32_t* n) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  if (n) *n = c.n_matched;
  if (!accept) return 0;
  if (c.n_matched > cap) return set_error(t->b->ctx, CUIMG_E_CAPACITY, "get_decisions: buffer too small");
  const level_res& R = t->b->lv[j];
  const size_t m = c.n_matched;
  hipStream_t st = t->b->ctx->stream;
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(accept, R.dec_accept, 4 * m, hipMemcpyDeviceToHost, st));
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(tr2, R.dec_tr, 8 * m, hipMemcpyDeviceToHost, st));
  CUIMG_HIP(t->b->ctx, hipMemcpyAsync(dist, R.dec_dist, 4 * m, hipMemcpyDeviceToHost, st));
  CUIMG_HIP(t->b->ctx, hipStreamSynchronize(st));
  return 0;
}

int cuimg_hip_tracker_matcher_images(cuimg_hip_tracker* t, int scale, cuimg_img* A, cuimg_img* B, cuimg_img* G) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, t->b->first[t->stream], &c));
  const level_res& R = t->b->lv[j];
  const int cur = (c.steps + 1) & 1, prev = cur ^ 1;
  const bool u8 = t->b->P0.profile == 0;
  if (A) *A = u8 ? R.u8[prev] : R.blur[prev];
  if (B) *B = u8 ? R.u8[cur] : R.blur[cur];
  if (G) {
    if (u8) std::memset(G, 0, sizeof(*G));  // the u8 profile computes it on the fly
    else *G = R.grad[prev];
  }
  return 0;
}

int cuimg_hip_tracker_work(cuimg_hip_tracker* t, int scale, uint64_t* live, uint64_t* accepted) {
  int j;
  CUIMG_TRY(job_of(t, scale, &j));
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  if (live) *live = c.live_total;
  if (accepted) *accepted = c.accepted_total;
  return 0;
}

int cuimg_hip_tracker_frame_count(cuimg_hip_tracker* t, int* out) {
  int j;
  CUIMG_TRY(job_of(t, 0, &j));
  if (!out) return CUIMG_E_INVALID;
  level_counters c;
  CUIMG_TRY(read_counters(t->b, j, &c));
  *out = (int)c.frame;
  return 0;
}

int cuimg_hip_particles_merge(cuimg_hip_ctx* ctx, cuimg_particle_f* parts, uint32_t n, const cuimg_img* sparse) {
  if (!ctx || (!parts && n) || !sparse || sparse->dtype != CUIMG_U32)
    return set_error(ctx, CUIMG_E_INVALID, "particles_merge: bad arguments");
  if (n > kMergeMaxCap)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "particles_merge: %u particles exceed %u", n, kMergeMaxCap);
  const uint32_t cap = std::max<uint32_t>(n, 1);
  void *buf = nullptr, *cnt = nullptr;
  CUIMG_HIP(ctx, hipMalloc(&buf, merge_scratch_bytes(cap)));
  if (hipMalloc(&cnt, sizeof(level_counters)) != hipSuccess) {
    (void)hipFree(buf);
    return set_error(ctx, CUIMG_E_NOMEM, "particles_merge: allocation failed");
  }
  level_counters c{};
  c.n = n;
  int rc = 0;
  if (hipMemcpyAsync(cnt, &c, sizeof c, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = CUIMG_E_NOMEM;
  if (!rc) {
    merge_args m = merge_scratch_carve(buf, cap);
    m.nr = sparse->nrows;
    m.nc = sparse->ncols;
    m.parts = parts;
    m.sparse = static_cast<const uint32_t*>(sparse->begin);
    m.sparse_pitch = sparse->pitch;
    m.cnt = static_cast<level_counters*>(cnt);
    rc = launch_merge_reset(ctx, m);
    if (!rc) rc = launch_merge(ctx, m);
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(buf);
  (void)hipFree(cnt);
  return rc;
}

int cuimg_hip_particle_mask(cuimg_hip_ctx* ctx, const cuimg_particle_f* parts, uint32_t n, const cuimg_img* mask) {
  if (!ctx || (!parts && n) || !mask || mask->dtype != CUIMG_U8 || !mask->begin || mask->nrows < 1 ||
      mask->ncols < 1)
    return set_error(ctx, CUIMG_E_INVALID, "particle_mask: bad arguments");
  uint8_t* m = static_cast<uint8_t*>(mask->begin);
  const int nr = mask->nrows, nc = mask->ncols, pitch = (int)mask->pitch;
  hipStream_t st = ctx->stream;
  CUIMG_TRY(launch(ctx, "mask_fill", [&] {
    mask_fill_kernel<<<std::min(nr, 1024), 256, 0, st>>>(m, pitch, nr, nc);
  }));
  if (n)
    CUIMG_TRY(launch(ctx, "particle_mask", [&] {
      mask_particles_kernel<<<std::min<uint32_t>(ceil_div((int)std::min<uint32_t>(n, 1u << 30), 256), 4096), 256,
                              0, st>>>(parts, n, m, pitch, nr, nc);
    }));
  return 0;
}

int cuimg_hip_fast_select(cuimg_hip_ctx* ctx, const cuimg_img* sal, const cuimg_img* sparse, int box, float th,
                          int32_t* pts, uint32_t cap, uint32_t* count) {
  if (!ctx || !sal || !sparse || !count || (!pts && cap) || sal->dtype != CUIMG_U8 || sparse->dtype != CUIMG_U32 ||
      !sal->begin || !sparse->begin || sal->nrows != sparse->nrows || sal->ncols != sparse->ncols ||
      sal->nrows < 1 || sal->ncols < 1)
    return set_error(ctx, CUIMG_E_INVALID, "fast_select: bad arguments");
  if (box < 1 || box > kMaxBox)
    return set_error(ctx, CUIMG_E_UNSUPPORTED, "fast_select: box %d not in 1..%d", box, kMaxBox);
  detect_args d{};
  d.nr = sal->nrows;
  d.nc = sal->ncols;
  d.box = box;
  d.nbr = (int)((float)d.nr / (float)box);
  d.nbc = (int)((float)d.nc / (float)box);
  d.fast_th = th;
  *count = 0;
  if (d.nbr < 1 || d.nbc < 1) return 0;
  d.sal = static_cast<const uint8_t*>(sal->begin);
  d.sal_pitch = (int)sal->pitch;
  d.sparse = static_cast<uint32_t*>(sparse->begin);
  d.sparse_pitch = (int)sparse->pitch;
  d.cap = cap;
  d.nblocks = 0;  // the scan's particle-block pass is empty
  const size_t nboxes = (size_t)d.nbr * d.nbc, nsub = (size_t)d.nbr * box;
  const size_t off_sub = (nboxes + 255) & ~(size_t)255, off_cnt = off_sub + ((nsub * 4 + 255) & ~(size_t)255);
  void* buf = nullptr;
  CUIMG_HIP(ctx, hipMalloc(&buf, off_cnt + sizeof(level_counters)));
  char* base = static_cast<char*>(buf);
  d.boxcode = reinterpret_cast<uint8_t*>(base);
  d.subcnt = reinterpret_cast<uint32_t*>(base + off_sub);
  d.cnt = reinterpret_cast<level_counters*>(base + off_cnt);
  hipStream_t st = ctx->stream;
  const hipError_t e = hipMemsetAsync(d.cnt, 0, sizeof(level_counters), st);
  int rc = e == hipSuccess ? 0 : hip_error(ctx, e, "hipMemsetAsync", __FILE__, __LINE__);
  if (!rc) rc = launch(ctx, "select_rows", [&] { select_rows_kernel<<<d.nbr, 256, 0, st>>>(d); });
  if (!rc) rc = launch(ctx, "select_scan", [&] { select_scan_kernel<<<1, 1024, 0, st>>>(d); });
  if (!rc && cap)
    rc = launch(ctx, "select_points", [&] { select_points_kernel<<<ceil_div((int)nsub, 4), 256, 0, st>>>(d, pts, cap); });
  level_counters c{};
  if (!rc) {
    const hipError_t ec = hipMemcpyAsync(&c, d.cnt, sizeof c, hipMemcpyDeviceToHost, st);
    if (ec != hipSuccess) rc = hip_error(ctx, ec, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
  if (!rc) *count = c.n_new;
  return rc;
}

int cuimg_hip_particles_add(cuimg_hip_ctx* ctx, cuimg_particle_f* parts, uint32_t n, uint32_t cap,
                            const int32_t* pts, uint32_t npts, const cuimg_img* sparse, uint32_t* n_out) {
  if (!ctx || !parts || !n_out || (!pts && npts) || !sparse || sparse->dtype != CUIMG_U32 || !sparse->begin ||
      n > cap || cap > kMaxParticles)
    return set_error(ctx, CUIMG_E_INVALID, "particles_add: bad arguments");
  const uint32_t m = std::min(npts, cap - n);
  *n_out = n;
  if (m) {
    hipStream_t st = ctx->stream;
    CUIMG_TRY(launch(ctx, "add_points", [&] {
      add_points_kernel<<<std::min(ceil_div((int)m, 256), 4096), 256, 0, st>>>(parts, n, m, pts);
    }));
    CUIMG_TRY(launch_claims(ctx, parts, n, n + m, 0, sparse));
  }
  *n_out = n + m;
  return 0;
}

int cuimg_hip_particles_compact(cuimg_hip_ctx* ctx, const cuimg_particle_f* parts, uint32_t n,
                                cuimg_particle_f* out, uint32_t* matches, const cuimg_img* sparse, uint32_t* n_out) {
  if (!ctx || !n_out || !sparse || sparse->dtype != CUIMG_U32 || !sparse->begin ||
      (n && (!parts || !out || !matches)) || n > kMaxParticles)
    return set_error(ctx, CUIMG_E_INVALID, "particles_compact: bad arguments");
  const uintptr_t a0 = (uintptr_t)parts, a1 = a0 + (uintptr_t)n * sizeof(cuimg_particle_f);
  const uintptr_t b0 = (uintptr_t)out, b1 = b0 + (uintptr_t)n * sizeof(cuimg_particle_f);
  if (n && a0 < b1 && b0 < a1) return set_error(ctx, CUIMG_E_INVALID, "particles_compact: out overlaps the input");
  *n_out = 0;
  if (!n) return 0;
  CUIMG_TRY(launch_claims(ctx, parts, 0, n, 1, sparse));
  detect_args d{};
  d.nr = sparse->nrows;
  d.nc = sparse->ncols;
  d.sparse = static_cast<uint32_t*>(sparse->begin);
  d.sparse_pitch = (int)sparse->pitch;
  d.parts = const_cast<cuimg_particle_f*>(parts);  // read only by the count and scatter passes
  d.parts_out = out;
  d.matches = matches;
  d.cap = n;
  d.nblocks = ceil_div((int)n, kCompactBlock);  // nbr = 0: the scan's sub-row pass is empty
  const size_t off_blk = ((size_t)n + 255) & ~(size_t)255;
  const size_t off_cnt = off_blk + (((size_t)d.nblocks * 4 + 255) & ~(size_t)255);
  void* buf = nullptr;
  CUIMG_HIP(ctx, hipMalloc(&buf, off_cnt + sizeof(level_counters)));
  char* base = static_cast<char*>(buf);
  d.flags = reinterpret_cast<uint8_t*>(base);
  d.blockcnt = reinterpret_cast<uint32_t*>(base + off_blk);
  d.cnt = reinterpret_cast<level_counters*>(base + off_cnt);
  hipStream_t st = ctx->stream;
  level_counters c{};
  c.n = n;
  int rc = 0;
  const hipError_t e = hipMemcpyAsync(d.cnt, &c, sizeof c, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) rc = hip_error(ctx, e, "hipMemcpyAsync", __FILE__, __LINE__);
  if (!rc) rc = launch(ctx, "compact_count", [&] { compact_count_kernel<<<d.nblocks, 256, 0, st>>>(d); });
  if (!rc) rc = launch(ctx, "compact_scan", [&] { compact_scan_kernel<<<1, 1024, 0, st>>>(d); });
  if (!rc) rc = launch(ctx, "compact_scatter", [&] { compact_scatter_kernel<<<d.nblocks, 256, 0, st>>>(d); });
  if (!rc) {
    const hipError_t ec = hipMemcpyAsync(&c, d.cnt, sizeof c, hipMemcpyDeviceToHost, st);
    if (ec != hipSuccess) rc = hip_error(ctx, ec, "hipMemcpyAsync", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
  if (!rc) *n_out = c.n_live;
  return rc;
}

}  // extern "C"
