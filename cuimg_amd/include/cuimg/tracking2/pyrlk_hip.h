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
    fast_detector& set_box_size(unsigned b)
    {
      if (st_->t) throw std::runtime_error("cuimg: box size is fixed once the tracker has run");
      st_->P.box_size = (int)b;
      return *this;
    }
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
      typedef MATCHER matcher;

      pyrlk_hip(pyrlk_hip_state* st = nullptr, int s = 0) : st_(st), s_(s), detector_(st, s) {}

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
      detector_t detector_;
    };
  }
}

#endif
