// The reference sample's call pattern (samples/tracking/main.cc:59-148) on the
// GPU strategy, without OpenCV: frames come from the synthetic generator.
//
//   tracking_sample nscales detector_threshold nframes rows cols
//
// For every frame it prints the particle count of each scale and the number
// of trajectories kept in sync with scale 0 through sync_attributes, then
// checks that the C++ facade and a tracker driven through the raw C ABI with
// the same settings end in the same state.  A second facade tracker is driven
// through update_input() + run(), with a smaller particle domain at scale 0
// and clear() of scale 1 half way; its twin on the C ABI gets the same calls.
// input() and get_flow_at() are checked at the end.
#include <cmath>
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

  // update_input() + run(), a particle domain and a cleared scale
  T1 tr2(domain, NSCALES);
  tr2.strategy().set_detector_frequency(1);
  tr2.strategy().set_filtering_frequency(1);
  for (int s = 0; s < NSCALES; s++) tr2.scale(s).strategy().detector().set_n(9).set_fast_threshold(th);
  const box2d pd(i_int2(rows / 5, cols / 6), i_int2(rows - 1 - rows / 4, cols - 1 - cols / 5));
  tr2.strategy().set_particle_domain(pd);
  cuimg_hip_tracker* raw2 = nullptr;
  ctx.check(cuimg_hip_tracker_create(ctx.get(), rows, cols, &P, &raw2), "create");
  cuimg_scale_params sp0;
  ctx.check(cuimg_hip_tracker_get_scale_params(raw2, 0, &sp0), "get_scale_params");
  sp0.pd_r1 = pd.p1().x; sp0.pd_c1 = pd.p1().y; sp0.pd_r2 = pd.p2().x; sp0.pd_c2 = pd.p2().y;
  ctx.check(cuimg_hip_tracker_set_scale_params(raw2, 0, &sp0), "set_scale_params");

  std::vector<std::vector<trajectory> > trajectories(NSCALES);
  for (int f = 0; f < nframes; f++)
  {
    for (int r = 0; r < rows; r++)
      std::memcpy(frame_gl.row(r), &video[((size_t)f * rows + r) * cols], cols);
    tr1.run(frame_gl);
    ctx.check(cuimg_hip_tracker_run_host(raw, &video[(size_t)f * rows * cols], cols), "run_host");
    if (f == nframes / 2 && NSCALES > 1)
    {
      tr2.scale(1).strategy().clear();
      ctx.check(cuimg_hip_tracker_clear_scale(raw2, 1), "clear_scale");
    }
    tr2.update_input(frame_gl);
    tr2.run();
    ctx.check(cuimg_hip_tracker_update_input_host(raw2, &video[(size_t)f * rows * cols], cols), "update_input");
    ctx.check(cuimg_hip_tracker_step(raw2), "step");
    for (const auto& q : tr2.pset().dense_particles())
      if (q.age > 1 && !pd.has(i_int2((int)q.pos.x, (int)q.pos.y)))
      {
        std::printf("FAIL: a tracked particle outside the particle domain\n");
        return 1;
      }

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
  T1* facades[2] = {&tr1, &tr2};
  cuimg_hip_tracker* raws[2] = {raw, raw2};
  for (int k = 0; k < 2; k++)
    for (int s = 0; s < NSCALES; s++)
    {
      const auto& a = facades[k]->scale(s).pset().dense_particles();
      uint32_t n = 0;
      ctx.check(cuimg_hip_tracker_get_particles(raws[k], s, nullptr, 0, &n), "get");
      std::vector<cuimg_particle_f> b(n);
      if (n) ctx.check(cuimg_hip_tracker_get_particles(raws[k], s, b.data(), n, &n), "get");
      if (a.size() != n || (n && std::memcmp(a.data(), b.data(), n * sizeof(cuimg_particle_f))))
      {
        std::printf("FAIL: facade %d and C ABI differ at scale %d (%zu vs %u)\n", k + 1, s, a.size(), n);
        return 1;
      }
    }
  {
    // the particle domain changed the result
    uint32_t n1 = 0, n2 = 0;
    ctx.check(cuimg_hip_tracker_get_particles(raw, 0, nullptr, 0, &n1), "get");
    ctx.check(cuimg_hip_tracker_get_particles(raw2, 0, nullptr, 0, &n2), "get");
    if (n1 == n2) { std::printf("FAIL: the particle domain had no effect\n"); return 1; }
  }

  // input(): the last frame at scale 0, clamp-filled border
  {
    const host_image2d<gl8u>& in = tr2.input();
    const uint8_t* last = &video[(size_t)(nframes - 1) * rows * cols];
    for (int r = -in.border(); r < rows + in.border(); r++)
      for (int c = -in.border(); c < cols + in.border(); c++)
      {
        const int rr = r < 0 ? 0 : (r >= rows ? rows - 1 : r), cc = c < 0 ? 0 : (c >= cols ? cols - 1 : c);
        if ((uint8_t)in(i_int2(r, c)) != last[(size_t)rr * cols + cc])
        {
          std::printf("FAIL: input() differs at (%d, %d)\n", r, c);
          return 1;
        }
      }
  }

  // get_flow_at: a particle of age 2 was detected at a pixel one frame step
  // ago and matched once, with the prediction a new particle gets; the flow at
  // that pixel between the last two frames is its speed, bit for bit
  {
    int checked = 0;
    const auto& ps = tr1.pset().dense_particles();
    for (size_t i = 0; i < ps.size() && checked < 64; i++)
    {
      if (ps[i].age != 2) continue;
      const i_float2 born(std::nearbyint(ps[i].pos.x - ps[i].speed.x), std::nearbyint(ps[i].pos.y - ps[i].speed.y));
      if (born.x + ps[i].speed.x != ps[i].pos.x || born.y + ps[i].speed.y != ps[i].pos.y) continue;
      const auto tr = tr1.strategy().get_flow_at(born);
      if (tr.x != ps[i].speed.x || tr.y != ps[i].speed.y)
      {
        std::printf("FAIL: get_flow_at(%g, %g) = (%g, %g), the frame step moved it by (%g, %g)\n", born.x, born.y,
                    tr.x, tr.y, ps[i].speed.x, ps[i].speed.y);
        return 1;
      }
      checked++;
    }
    if (!checked) { std::printf("FAIL: no particle of age 2 to check get_flow_at on\n"); return 1; }
    const auto far = tr1.strategy().get_flow_at(i_float2(-50.f, -50.f));
    if (far.x != 0.f || far.y != 0.f) { std::printf("FAIL: get_flow_at outside the domain\n"); return 1; }
    std::printf("get_flow_at: %d frame-step matches reproduced\n", checked);
  }
  cuimg_hip_tracker_destroy(raw);
  cuimg_hip_tracker_destroy(raw2);
  std::printf("OK\n");
  return 0;
}
