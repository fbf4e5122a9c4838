// The reference sample's call pattern (samples/tracking/main.cc:59-148) on the
// GPU strategy, without OpenCV: frames come from the synthetic generator.
//
//   tracking_sample nscales detector_threshold nframes rows cols
//
// For every frame it prints the particle count of each scale and the number
// of trajectories kept in sync with scale 0 through sync_attributes, then
// checks that the C++ facade and a tracker driven through the raw C ABI with
// the same settings end in the same state.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <cuimg/tracking2/tracker.h>

using namespace cuimg;

struct trajectory
{
  std::vector<i_float2> history;
  bool alive = true;
};

int main(int argc, char* argv[])
{
  if (argc < 6)
  {
    std::printf("usage: tracking_sample nscales detector_threshold nframes rows cols\n");
    return 2;
  }
  const int NSCALES = std::atoi(argv[1]);
  const int th = std::atoi(argv[2]);
  const int nframes = std::atoi(argv[3]);
  const int rows = std::atoi(argv[4]), cols = std::atoi(argv[5]);

  std::vector<uint8_t> video((size_t)nframes * rows * cols);
  if (cuimg_synth_video(0xC0FFEE01ull, rows, cols, nframes, 1, video.data())) return 3;

  obox2d domain(rows, cols);
  host_image2d<gl8u> frame_gl(domain);

  typedef tracker<tracking_strategies::pyrlk_hip<fast_detector<hip_gpu>, blurred_lk_match_7x7> > T1;
  T1 tr1(domain, NSCALES);
  tr1.strategy().set_detector_frequency(1);
  tr1.strategy().set_filtering_frequency(1);
  for (int s = 0; s < NSCALES; s++) tr1.scale(s).strategy().detector().set_n(9).set_fast_threshold(th);

  // the same settings through the C ABI
  hip_context& ctx = hip_context::default_context();
  cuimg_tracker_params P{};
  P.nscales = NSCALES; P.profile = 1; P.lk_ws = 7; P.lk_max_iters = 5; P.lk_eps = 0.1f; P.lk_top_n = 49;
  P.fast_n = 9; P.fast_th = (float)th; P.box_size = 3; P.detector_frequency = 1; P.filtering_frequency = 1;
  P.k = 300; P.with_merge = 1; P.blur_sigma = 1; P.blur_K = 3;
  cuimg_hip_tracker* raw = nullptr;
  ctx.check(cuimg_hip_tracker_create(ctx.get(), rows, cols, &P, &raw), "create");

  std::vector<std::vector<trajectory> > trajectories(NSCALES);
  for (int f = 0; f < nframes; f++)
  {
    for (int r = 0; r < rows; r++)
      std::memcpy(frame_gl.row(r), &video[((size_t)f * rows + r) * cols], cols);
    tr1.run(frame_gl);
    ctx.check(cuimg_hip_tracker_run_host(raw, &video[(size_t)f * rows * cols], cols), "run_host");

    for (int s = 0; s < NSCALES; s++)
    {
      tr1.scale(s).pset().sync_attributes(trajectories[s], trajectory());
      const auto& parts = tr1.scale(s).pset().dense_particles();
      if (trajectories[s].size() != parts.size())
      {
        std::printf("FAIL: %zu trajectories for %zu particles at scale %d\n", trajectories[s].size(), parts.size(), s);
        return 1;
      }
      for (size_t i = 0; i < parts.size(); i++)
      {
        auto& h = trajectories[s][i].history;
        if (parts[i].age == 1 && !h.empty()) { std::printf("FAIL: new particle inherits a history\n"); return 1; }
        h.push_back(parts[i].pos);
      }
    }
    std::printf("frame %d:", f);
    for (int s = 0; s < NSCALES; s++) std::printf(" %d", tr1.scale(s).pset().size());
    std::printf(" (total %u)\n", tr1.nparticles());
  }

  // facade and raw C ABI: identical device computations, identical state
  for (int s = 0; s < NSCALES; s++)
  {
    const auto& a = tr1.scale(s).pset().dense_particles();
    uint32_t n = 0;
    ctx.check(cuimg_hip_tracker_get_particles(raw, s, nullptr, 0, &n), "get");
    std::vector<cuimg_particle_f> b(n);
    if (n) ctx.check(cuimg_hip_tracker_get_particles(raw, s, b.data(), n, &n), "get");
    if (a.size() != n || (n && std::memcmp(a.data(), b.data(), n * sizeof(cuimg_particle_f))))
    {
      std::printf("FAIL: facade and C ABI differ at scale %d (%zu vs %u)\n", s, a.size(), n);
      return 1;
    }
  }
  cuimg_hip_tracker_destroy(raw);
  std::printf("OK\n");
  return 0;
}
