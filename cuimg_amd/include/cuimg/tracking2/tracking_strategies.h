// The bc2s tracking strategy on the GPU: the policy type that plugs into
// tracker<F> in place of tracking_strategies::generic_strategy<bc2s_feature<A>,
// fast_detector<A>, particle_container<...>, image> (tracking2/tracking_strategies.h:28-111),
// i.e. the reference sample's bc2s_fast_gradient_cpu (samples/tracking/main.cc:101).
// Setters keep the reference's names and propagation; the device tracker runs
// every scale's frame step.
#ifndef CUIMG_TRACKING2_TRACKING_STRATEGIES_H_
# define CUIMG_TRACKING2_TRACKING_STRATEGIES_H_

# include <algorithm>

# include <cuimg/tracking2/pyrlk_hip.h>

namespace cuimg
{
  // Matcher tag of the bc2s strategy: the device tracker's profile 2 (the LK
  // fields are unused there).
  struct bc2s_gradient_descent_match
  {
    typedef translation_transform transform;
    static const int profile = 2, ws = 7, max_iters = 5, top_n = 49;
  };

  // bc2s_feature<A>'s images of one scale (s1(), s2(), bc2s_feature.h:76-77),
  // copied from the device on demand, borders included.
  class bc2s_feature_host
  {
  public:
    typedef host_image2d<gl8u> I;
    bc2s_feature_host(pyrlk_hip_state* st, int s) : st_(st), s_(s) {}
    const I& s1() const { load(); return s1_; }
    const I& s2() const { load(); return s2_; }
    int border_needed() const { return 6; }

  private:
    void load() const
    {
      if (!st_->t) return;
      cuimg_img a, b;
      st_->ctx->check(cuimg_hip_tracker_bc2s_images(st_->t, s_, &a, &b, nullptr), "bc2s images");
      get(a, s1_);
      get(b, s2_);
    }
    void get(const cuimg_img& v, I& out) const
    {
      const int bd = v.border;
      if (out.nrows() != v.nrows || out.ncols() != v.ncols || out.border() != bd) out = I(obox2d(v.nrows, v.ncols), bd);
      cuimg_img all = v;
      all.begin = static_cast<uint8_t*>(v.begin) - (ptrdiff_t)bd * v.pitch - bd;
      all.nrows = v.nrows + 2 * bd;
      all.ncols = v.ncols + 2 * bd;
      all.border = 0;
      uint8_t* host = reinterpret_cast<uint8_t*>(out.begin()) - (ptrdiff_t)bd * out.pitch() - bd;
      st_->ctx->check(cuimg_hip_image_download(st_->ctx->get(), &all, host, out.pitch()), "bc2s images");
    }
    pyrlk_hip_state* st_;
    int s_;
    mutable I s1_, s2_;
  };

  namespace tracking_strategies
  {
    template <typename DETECTOR>
    class generic_hip
    {
    public:
      typedef translation_transform transform;
      typedef DETECTOR detector_t;
      typedef bc2s_feature_host feature_t;
      typedef particle_container_view_t<particle> particles_type;
      typedef particle particle_type;
      typedef hip_gpu architecture;
      typedef host_image2d<gl8u> input;
      typedef host_image2d<i_float2> flow_t;
      typedef bc2s_gradient_descent_match matcher;

      generic_hip(pyrlk_hip_state* st = nullptr, int s = 0) : st_(st), s_(s), detector_(st, s), feature_(st, s) {}

      void init() {}
      void set_upper(generic_hip* u) { upper_ = u; }
      generic_hip* upper() { return upper_; }

      // generic_strategy::update (tracking_strategies.hpp:155-198), driven from
      // the finest scale (the device step advances every scale at once).
      void update(const input& in, particles_type&)
      {
        if (s_ != 0 || !st_->t) return;
        st_->ctx->check(cuimg_hip_tracker_update_input_host(st_->t, reinterpret_cast<const uint8_t*>(in.begin()),
                                                            in.pitch()), "update_input");
        st_->ctx->check(cuimg_hip_tracker_step(st_->t), "step");
      }

      // generic_strategy::clear (tracking_strategies.hpp:755-760): this
      // scale's frame counter restarts at 0.
      void clear()
      {
        if (st_->t) st_->ctx->check(cuimg_hip_tracker_clear_scale(st_->t, s_), "clear");
      }

      // The setters propagate to the coarser scales like the reference's
      // (tracking_strategies.hpp:63-91); set_with_merge does not.
      generic_hip& set_detector_frequency(unsigned nframe)
      {
        for (size_t s = s_; s < st_->sp.size(); s++) { st_->sp[s].detector_frequency = (int)nframe; st_->push((int)s); }
        return *this;
      }
      generic_hip& set_filtering_frequency(unsigned nframe)
      {
        for (size_t s = s_; s < st_->sp.size(); s++) { st_->sp[s].filtering_frequency = (int)nframe; st_->push((int)s); }
        return *this;
      }
      generic_hip& set_k(int k)
      {
        for (size_t s = s_; s < st_->sp.size(); s++) { st_->sp[s].k = k; st_->push((int)s); }
        return *this;
      }
      void set_with_merge(bool b) { st_->sp[s_].with_merge = b ? 1 : 0; st_->push(s_); }

      detector_t& detector() { return detector_; }
      feature_t& feature() { return feature_; }
      int border_needed() const { return std::max(feature_.border_needed(), detector_.border_needed()); }

      // flow() (tracking_strategies.h:82): one translation per 8x8 cell of the
      // last frame, (-500, -500) where there is none.
      const flow_t& flow() const
      {
        if (!st_->t) return flow_;
        cuimg_img f;
        st_->ctx->check(cuimg_hip_tracker_flow(st_->t, s_, &f, nullptr), "flow");
        if (flow_.nrows() != f.nrows || flow_.ncols() != f.ncols) flow_ = flow_t(obox2d(f.nrows, f.ncols));
        st_->ctx->check(cuimg_hip_image_download(st_->ctx->get(), &f, flow_.begin(), flow_.pitch()), "flow");
        return flow_;
      }

      // estimate_camera_motion sets both to zero (tracking_strategies.hpp:717-723).
      i_short2 camera_motion() const { return i_short2(0, 0); }
      i_short2 prev_camera_motion() const { return i_short2(0, 0); }

    private:
      pyrlk_hip_state* st_;
      int s_;
      generic_hip* upper_ = nullptr;
      detector_t detector_;
      feature_t feature_;
      mutable flow_t flow_;
    };

    // The reference sample's strategy (samples/tracking/main.cc:101) and its
    // GPU twin (tracking_strategies.h:184-199): one strategy, on the GPU.
    typedef generic_hip<fast_detector<hip_gpu> > bc2s_fast_gradient_gpu;
    typedef bc2s_fast_gradient_gpu bc2s_fast_gradient_cpu;
  }
}

#endif
