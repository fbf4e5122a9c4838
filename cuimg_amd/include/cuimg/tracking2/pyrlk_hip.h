// Pyramid-LK tracking strategy on the GPU: the policy type that plugs into
// tracker<F> in place of tracking_strategies::pyrlk_cpu<fast_detector<cpu>,
// iterative_lk_match_5x5> (tracking2/pyrlk_tracker.h:16-84).  The setters keep
// the reference's names and per-scale meaning; they reach the device tracker
// through the C ABI.
#ifndef CUIMG_TRACKING2_PYRLK_HIP_H_
# define CUIMG_TRACKING2_PYRLK_HIP_H_

# include <functional>
# include <stdexcept>
# include <vector>

# include <cuimg/cpu/host_image2d.h>
# include <cuimg/gl.h>
# include <cuimg/hip/context.h>
# include <cuimg/obox2d.h>
# include <cuimg/tracking2/particle_container.h>

namespace cuimg
{
  struct hip_gpu {};   // architecture tag of the device path

  // Matcher tags: the reference's defaults, and the blurred-frame variant.
  struct iterative_lk_match_5x5
  {
    typedef translation_transform transform;
    static const int profile = 0, ws = 11, max_iters = 30, top_n = 25;
  };
  struct blurred_lk_match_7x7
  {
    typedef translation_transform transform;
    static const int profile = 1, ws = 7, max_iters = 5, top_n = 49;
  };

  // Per-scale settings shared by the strategy / detector facades of a tracker.
  struct pyrlk_hip_state
  {
    hip_context* ctx = nullptr;
    cuimg_tracker_params P{};
    std::vector<cuimg_scale_params> sp;
    std::vector<obox2d> domains;  // of every scale
    cuimg_hip_tracker* t = nullptr;

    void push(int s)
    {
      if (t) ctx->check(cuimg_hip_tracker_set_scale_params(t, s, &sp[s]), "set_scale_params");
    }
  };

  template <typename A>
  class fast_detector
  {
  public:
    fast_detector(pyrlk_hip_state* st = nullptr, int s = 0) : st_(st), s_(s) {}
    fast_detector& set_fast_threshold(float f) { st_->sp[s_].fast_th = f; st_->push(s_); return *this; }
    fast_detector& set_n(unsigned n) { st_->sp[s_].fast_n = (int)n; st_->push(s_); return *this; }
    fast_detector& set_box_size(unsigned b) { st_->sp[s_].box_size = (int)b; st_->push(s_); return *this; }
    int border_needed() const { return 3; }
  private:
    pyrlk_hip_state* st_;
    int s_;
  };

  namespace tracking_strategies
  {
    template <typename DETECTOR, typename MATCHER>
    class pyrlk_hip
    {
    public:
      typedef typename MATCHER::transform transform;
      typedef DETECTOR detector_t;
      typedef particle_container_view particles_type;
      typedef particle_f<transform> particle_type;
      typedef hip_gpu architecture;
      typedef host_image2d<gl8u> input;
      typedef MATCHER matcher;

      pyrlk_hip(pyrlk_hip_state* st = nullptr, int s = 0) : st_(st), s_(s), detector_(st, s) {}

      void init() {}
      // The scales of one device tracker are linked already: the device step
      // runs the coarse-to-fine recursion itself.
      void set_upper(pyrlk_hip* u) { upper_ = u; }
      pyrlk_hip* upper() { return upper_; }

      // One frame step on `in` (pyrlk_cpu::update, pyrlk_tracker.hpp:831-875).
      // The device step advances every scale of the tracker at once, so it is
      // driven from the finest scale; on a coarser scale this does nothing
      // (tracker<F>::run reaches every scale through the finest one).
      void update(const input& in, particles_type&)
      {
        if (s_ != 0 || !st_->t) return;
        st_->ctx->check(cuimg_hip_tracker_update_input_host(st_->t, reinterpret_cast<const uint8_t*>(in.begin()),
                                                            in.pitch()), "update_input");
        st_->ctx->check(cuimg_hip_tracker_step(st_->t), "step");
      }

      // pyrlk_cpu::clear: this scale's frame counter restarts at 0.
      void clear()
      {
        if (st_->t) st_->ctx->check(cuimg_hip_tracker_clear_scale(st_->t, s_), "clear");
      }

      // pyrlk_cpu::set_particle_domain (pyrlk_tracker.hpp:894-897).
      void set_particle_domain(const box2d& d)
      {
        cuimg_scale_params& p = st_->sp[s_];
        p.pd_r1 = d.p1().x;
        p.pd_c1 = d.p1().y;
        p.pd_r2 = d.p2().x;
        p.pd_c2 = d.p2().y;
        st_->push(s_);
      }

      // pyrlk_cpu::get_flow_at (pyrlk_tracker.hpp:1025-1063): the match of a
      // new (age 1) particle at p between the last two frames of this scale,
      // predicted through the coarser scales; identity when the prediction or
      // the match leaves the domain (or before two frames were seen).
      transform get_flow_at(const i_float2& p)
      {
        if (!st_->t) return transform::identity();
        hip_context& ctx = *st_->ctx;
        const int L = st_->P.nscales;
        std::vector<cuimg_img> A(L), B(L);
        for (int k = 0; k < L; k++)
          ctx.check(cuimg_hip_tracker_matcher_images(st_->t, k, &A[k], &B[k], nullptr), "matcher_images");
        cuimg_particle_f part{0.f, 0.f, p.x, p.y, 1, 0};
        void* d = nullptr;
        ctx.check(cuimg_hip_malloc(ctx.get(), 64, &d), "malloc");
        char* dc = static_cast<char*>(d);
        float tr[2] = {0.f, 0.f};
        int rc = cuimg_hip_memcpy_h2d(ctx.get(), dc, &part, sizeof part);
        const cuimg_lk_params prm{st_->P.lk_ws, st_->P.lk_max_iters, st_->P.lk_eps, st_->P.lk_top_n, st_->sp[s_].k,
                                  st_->P.profile == 0 ? 1 : 0};
        if (!rc)
          rc = cuimg_hip_lk_track(ctx.get(), A.data(), B.data(), L, s_, reinterpret_cast<cuimg_particle_f*>(dc), 1,
                                  &prm, nullptr, reinterpret_cast<int32_t*>(dc + 32), reinterpret_cast<float*>(dc + 40),
                                  reinterpret_cast<float*>(dc + 48));
        if (!rc) rc = cuimg_hip_memcpy_d2h(ctx.get(), tr, dc + 40, sizeof tr);
        cuimg_hip_free(ctx.get(), d);
        ctx.check(rc, "get_flow_at");
        const obox2d& dom = st_->domains[s_];
        const i_int2 m((int)(p.x + tr[0]), (int)(p.y + tr[1]));
        return dom.has(m) ? transform(tr[0], tr[1]) : transform::identity();
      }

      // set_detector_frequency / set_filtering_frequency propagate to the
      // coarser scales (pyrlk_tracker.hpp:802-821); set_k / set_with_merge do not.
      pyrlk_hip& set_detector_frequency(unsigned nframe)
      {
        for (size_t s = s_; s < st_->sp.size(); s++) { st_->sp[s].detector_frequency = (int)nframe; st_->push((int)s); }
        return *this;
      }
      pyrlk_hip& set_filtering_frequency(unsigned nframe)
      {
        for (size_t s = s_; s < st_->sp.size(); s++) { st_->sp[s].filtering_frequency = (int)nframe; st_->push((int)s); }
        return *this;
      }
      pyrlk_hip& set_k(int k) { st_->sp[s_].k = k; st_->push(s_); return *this; }
      void set_with_merge(bool b) { st_->sp[s_].with_merge = b ? 1 : 0; st_->push(s_); }
      detector_t& detector() { return detector_; }
      int border_needed() const { return detector_.border_needed(); }

    private:
      pyrlk_hip_state* st_;
      int s_;
      pyrlk_hip* upper_ = nullptr;
      detector_t detector_;
    };
  }
}

#endif
