// tracker<F> for the GPU strategy (tracking2/tracker.h:10-47, .hpp:12-112):
// same construction, run(in), update_input(in) + run(), input(), scale(s),
// pset(), strategy(), clear() and nparticles().  All scales live in one device tracker; scale(s) returns the
// facade of scale s.  The device state is created at the first run(), after
// the setters the caller applies to the freshly built tracker.
#ifndef CUIMG_TRACKING2_TRACKER_H_
# define CUIMG_TRACKING2_TRACKER_H_

# include <memory>
# include <stdexcept>
# include <vector>

# include <cuimg/cpu/host_image2d.h>
# include <cuimg/gl.h>
# include <cuimg/gpu/device_image2d.h>
# include <cuimg/tracking2/pyrlk_hip.h>
# include <cuimg/tracking2/tracking_strategies.h>

namespace cuimg
{
  template <typename F>
  class tracker
  {
  public:
    typedef host_image2d<gl8u> I;
    typedef typename F::particles_type particles_type;
    typedef typename F::particle_type particle_type;

    tracker(const obox2d& d, int nscales, hip_context& ctx = hip_context::default_context())
      : owner_(new state()), st_(owner_.get()), s_(0), domain_(d)
    {
      if (nscales < 1) throw std::runtime_error("cuimg: nscales must be >= 1");
      state& S = *st_;
      S.h.ctx = &ctx;
      typedef typename F::matcher M;
      cuimg_tracker_params& P = S.h.P;
      P.nscales = nscales;
      P.profile = M::profile;
      P.lk_ws = M::ws;
      P.lk_max_iters = M::max_iters;
      P.lk_eps = 0.1f;
      P.lk_top_n = M::top_n;
      P.fast_n = 9;          // fast_detector ctor (fast_detector.hpp:160-168)
      P.fast_th = 0.f;       // the reference leaves it uninitialised: set it
      P.box_size = 3;
      P.detector_frequency = 1;  // pyrlk_cpu ctor (pyrlk_tracker.hpp:771-793)
      P.filtering_frequency = 1;
      P.k = 300;
      P.with_merge = 1;
      P.blur_sigma = 1;
      P.blur_K = 3;
      obox2d ds = d;
      for (int s = 0; s < nscales; s++)
      {
        S.domains.push_back(ds);
        // particle domain: domain - border(3) (pyrlk_tracker.hpp:792)
        S.h.sp.push_back(cuimg_scale_params{P.fast_n, P.fast_th, P.detector_frequency, P.filtering_frequency, P.k,
                                            P.with_merge, P.box_size, 3, 3, ds.nrows() - 4, ds.ncols() - 4});
        ds = ds / 2;
      }
      S.h.domains = S.domains;
      S.scales.push_back(this);
      for (int s = 1; s < nscales; s++)
        S.owned.emplace_back(new tracker(st_, s, S.domains[s])), S.scales.push_back(S.owned.back().get());
      // tracker<F> ctor (tracker.hpp:12-26): each strategy knows its coarser
      // neighbour, then init()
      for (int s = 0; s + 1 < nscales; s++) S.scales[s]->strategy_.set_upper(&S.scales[s + 1]->strategy_);
      for (int s = 0; s < nscales; s++) S.scales[s]->strategy_.init();
    }

    tracker(const tracker&) = delete;
    tracker& operator=(const tracker&) = delete;

    ~tracker()
    {
      if (owner_ && owner_->h.t) cuimg_hip_tracker_destroy(owner_->h.t);
    }

    // A host frame goes through the library's pinned double buffer: the call
    // returns once the frame is staged, and its upload overlaps the previous
    // frame step.
    tracker& run(const I& in)
    {
      if (s_ != 0) throw std::runtime_error("cuimg: run() is called on the finest scale");
      ensure();
      st_->h.ctx->check(cuimg_hip_tracker_run_host(st_->h.t, reinterpret_cast<const uint8_t*>(in.begin()), in.pitch()),
                        "tracker run");
      return *this;
    }

    tracker& run(const device_image2d<gl8u>& in)
    {
      if (s_ != 0) throw std::runtime_error("cuimg: run() is called on the finest scale");
      ensure();
      st_->h.ctx->check(cuimg_hip_tracker_run(st_->h.t, &in.kernel()), "tracker run");
      return *this;
    }

    // tracker::update_input (tracker.hpp:53-69): the pyramid of the next
    // frame is built on the device; run() then steps on it (tracker.hpp:83-91;
    // without a new input it steps on the previous one again).
    tracker& update_input(const I& in)
    {
      if (s_ != 0) throw std::runtime_error("cuimg: update_input() is called on the finest scale");
      ensure();
      st_->h.ctx->check(cuimg_hip_tracker_update_input_host(st_->h.t, reinterpret_cast<const uint8_t*>(in.begin()),
                                                            in.pitch()), "update_input");
      return *this;
    }

    tracker& update_input(const device_image2d<gl8u>& in)
    {
      if (s_ != 0) throw std::runtime_error("cuimg: update_input() is called on the finest scale");
      ensure();
      st_->h.ctx->check(cuimg_hip_tracker_update_input(st_->h.t, &in.kernel()), "update_input");
      return *this;
    }

    tracker& run()
    {
      if (s_ != 0) throw std::runtime_error("cuimg: run() is called on the finest scale");
      ensure();
      st_->h.ctx->check(cuimg_hip_tracker_step(st_->h.t), "tracker step");
      return *this;
    }

    // input(): this scale's input level of the last frame (border 3,
    // clamp-filled), copied from the device.
    const I& input() const
    {
      if (!st_->h.t) return input_;
      cuimg_img v;
      st_->h.ctx->check(cuimg_hip_tracker_level(st_->h.t, s_, &v), "input");
      const int b = v.border;
      if (input_.nrows() != v.nrows || input_.ncols() != v.ncols || input_.border() != b)
        input_ = I(obox2d(v.nrows, v.ncols), b);
      cuimg_img all = v;  // domain plus border
      all.begin = static_cast<uint8_t*>(v.begin) - (ptrdiff_t)b * v.pitch - b;
      all.nrows = v.nrows + 2 * b;
      all.ncols = v.ncols + 2 * b;
      all.border = 0;
      uint8_t* host = reinterpret_cast<uint8_t*>(input_.begin()) - (ptrdiff_t)b * input_.pitch() - b;
      st_->h.ctx->check(cuimg_hip_image_download(st_->h.ctx->get(), &all, host, input_.pitch()), "input");
      return input_;
    }

    const obox2d& domain() const { return domain_; }
    tracker& scale(unsigned s) { return *st_->scales.at(s); }
    particles_type& pset() { return pset_; }
    F& strategy() { return strategy_; }

    void clear()
    {
      if (st_->h.t) st_->h.ctx->check(cuimg_hip_tracker_clear(st_->h.t), "clear");
    }

    unsigned nparticles() const
    {
      unsigned n = 0;
      for (size_t s = s_; s < st_->scales.size(); s++) n += st_->scales[s]->pset_.size();
      return n;
    }

    cuimg_hip_tracker* device_tracker() const { return st_->h.t; }

  private:
    struct state
    {
      pyrlk_hip_state h;
      std::vector<obox2d> domains;
      std::vector<tracker*> scales;
      std::vector<std::unique_ptr<tracker>> owned;
    };

    tracker(state* st, int s, const obox2d& d)
      : st_(st), s_(s), domain_(d), strategy_(&st->h, s),
        pset_([st]() { return st->h.t; }, [st]() { return st->h.ctx; }, s) {}

    void ensure()
    {
      pyrlk_hip_state& h = st_->h;
      if (h.t) return;
      h.ctx->check(cuimg_hip_tracker_create(h.ctx->get(), domain_.nrows(), domain_.ncols(), &h.P, &h.t),
                   "tracker create");
      for (size_t s = 0; s < h.sp.size(); s++) h.push((int)s);
    }

    mutable I input_;               // input() cache
    std::unique_ptr<state> owner_;  // the finest scale owns the shared state
    state* st_;
    int s_;
    obox2d domain_;
    F strategy_{&st_->h, s_};
    particles_type pset_{[st = st_]() { return st->h.t; }, [st = st_]() { return st->h.ctx; }, s_};
  };
}

#endif
