// The reference sample (samples/tracking/main.cc:59-148) with its own
// strategy, tracker<tracking_strategies::bc2s_fast_gradient_cpu>, and its
// settings (detector every 10 frames, filtering every frame, FAST n = 9),
// without OpenCV: frames come from the synthetic generator.
//
//   bc2s_sample nscales detector_threshold nframes rows cols
//
// Trajectories are kept in sync with every scale through sync_attributes and
// grow by the particle positions (update_trajectories).  At the end the
// facade's state (particles, descriptors) is compared with a tracker driven
// through the raw C ABI with the same settings, and the feature images and
// flow are read through the strategy.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <cuimg/tracking2/tracker.h>

using namespace cuimg;

struct trajectory
{
  std::vector<i_int2> history;
  bool alive = true;
};

template <typename V, typename P>
void update_trajectories(V& v, const P& pset)
{
  const auto& parts = pset.dense_particles();
  for (size_t i = 0; i < parts.size(); i++)
    if (parts[i].age > 0) v[i].history.push_back(i_int2(parts[i].pos.x, parts[i].pos.y));
    else v[i].history.clear();
}

int main(int argc, char* argv[])
{
  if (argc < 6)
  {
    std::printf("usage: bc2s_sample nscales detector_threshold nframes rows cols\n");
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

  // Tracker definition and settings, as the reference sample writes them
  typedef tracker<tracking_strategies::bc2s_fast_gradient_cpu> T1;
  T1 tr1(domain, NSCALES);
  tr1.strategy().set_detector_frequency(10);
  tr1.strategy().set_filtering_frequency(1);
  for (int s = 0; s < NSCALES; s++) tr1.scale(s).strategy().detector().set_n(9).set_fast_threshold(th);

  // the same settings through the C ABI
  hip_context& ctx = hip_context::default_context();
  cuimg_tracker_params P{};
  P.nscales = NSCALES; P.profile = 2; P.lk_ws = 7; P.lk_max_iters = 5; P.lk_eps = 0.1f; P.lk_top_n = 49;
  P.fast_n = 9; P.fast_th = (float)th; P.box_size = 3; P.detector_frequency = 10; P.filtering_frequency = 1;
  P.k = 300; P.with_merge = 1; P.blur_sigma = 1; P.blur_K = 3;
  cuimg_hip_tracker* raw = nullptr;
  ctx.check(cuimg_hip_tracker_create(ctx.get(), rows, cols, &P, &raw), "create");

  std::vector<std::vector<trajectory> > trajectories(NSCALES);
  size_t long_lived = 0;
  for (int f = 0; f < nframes; f++)
  {
    for (int r = 0; r < rows; r++)
      std::memcpy(frame_gl.row(r), &video[((size_t)f * rows + r) * cols], cols);
    tr1.run(frame_gl);
    ctx.check(cuimg_hip_tracker_run_host(raw, &video[(size_t)f * rows * cols], cols), "run_host");
    for (int s = 0; s < NSCALES; s++)
    {
      tr1.scale(s).pset().sync_attributes(trajectories[s], trajectory());
      update_trajectories(trajectories[s], tr1.scale(s).pset());
    }
    std::printf("frame %d:", f);
    for (int s = 0; s < NSCALES; s++) std::printf(" %d", tr1.scale(s).pset().size());
    std::printf(" (total %u)\n", tr1.nparticles());
  }
  for (const trajectory& t : trajectories[0]) long_lived += t.history.size() >= 3;
  if (!long_lived) { std::printf("FAIL: no trajectory of 3 frames\n"); return 1; }

  // an integer-position particle moves by integer steps along its history
  for (const trajectory& t : trajectories[0])
    for (size_t k = 1; k < t.history.size(); k++)
    {
      const i_int2 d = t.history[k] - t.history[k - 1];
      if (d.x < -60 || d.x > 60 || d.y < -60 || d.y > 60) { std::printf("FAIL: jump\n"); return 1; }
    }

  for (int s = 0; s < NSCALES; s++)
  {
    const auto& a = tr1.scale(s).pset().dense_particles();
    const auto& fa = tr1.scale(s).pset().features();
    uint32_t n = 0;
    ctx.check(cuimg_hip_tracker_get_particles(raw, s, nullptr, 0, &n), "get");
    std::vector<cuimg_particle_f> b(n);
    std::vector<uint8_t> fb(16 * (size_t)n);
    if (n)
    {
      ctx.check(cuimg_hip_tracker_get_particles(raw, s, b.data(), n, &n), "get");
      ctx.check(cuimg_hip_tracker_get_features(raw, s, fb.data(), n, &n), "features");
    }
    if (a.size() != n || fa.size() != n)
    {
      std::printf("FAIL: facade and C ABI differ in size at scale %d (%zu vs %u)\n", s, a.size(), n);
      return 1;
    }
    for (uint32_t i = 0; i < n; i++)
      if (a[i].pos.x != (short)b[i].pos_r || a[i].pos.y != (short)b[i].pos_c || a[i].age != b[i].age ||
          a[i].speed.x != b[i].speed_r || a[i].speed.y != b[i].speed_c ||
          std::memcmp(fa[i].distances, &fb[16 * (size_t)i], 16) || fa[i].weights[0] != 255)
      {
        std::printf("FAIL: facade and C ABI differ at scale %d particle %u\n", s, i);
        return 1;
      }
  }
  // the strategy's feature images and flow
  const auto& s1 = tr1.strategy().feature().s1();
  const auto& fl = tr1.strategy().flow();
  if (s1.nrows() != rows || s1.border() != 3 || fl.nrows() != (rows + 7) / 8 || fl.ncols() != (cols + 7) / 8)
  {
    std::printf("FAIL: feature / flow shapes\n");
    return 1;
  }
  std::printf("%zu trajectories of 3+ frames\n", long_lived);
  cuimg_hip_tracker_destroy(raw);
  std::printf("OK\n");
  return 0;
}
