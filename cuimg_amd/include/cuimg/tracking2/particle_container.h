// Particle container of one tracker scale, read from the device
// (tracking2/particle_container.h:13-225, .hpp:22-578).  The dense AoS layout
// particle_f<translation_transform> is the reference's: 20 bytes,
// {speed(row,col), pos(row,col), unsigned short age}.
#ifndef CUIMG_TRACKING2_PARTICLE_CONTAINER_H_
# define CUIMG_TRACKING2_PARTICLE_CONTAINER_H_

# include <functional>
# include <vector>

# include <cuimg/cpu/host_image2d.h>
# include <cuimg/hip/context.h>

namespace cuimg
{
  // translation_transform (tracking2/transformations.h:58-88)
  struct translation_transform : public i_float2
  {
    translation_transform() : i_float2(0.f, 0.f) {}
    translation_transform(float r, float c) : i_float2(r, c) {}
    translation_transform(const i_float2& v) : i_float2(v) {}
    static translation_transform identity() { return translation_transform(); }
    static translation_transform zero() { return translation_transform(); }
    translation_transform scale_transform(float) const { return translation_transform(2 * x, 2 * y); }
    i_float2 apply_transform(i_float2 p) const { return i_float2(p.x + x, p.y + y); }
  };

  template <typename T>
  struct particle_f
  {
    typedef i_float2 coords_type;
    typedef T transform;
    particle_f() : speed(T::identity()), pos(), age(0) {}
    T speed;
    i_float2 pos;
    unsigned short age;
  };
  static_assert(sizeof(particle_f<translation_transform>) == sizeof(cuimg_particle_f),
                "particle_f<translation_transform> is the 20-byte cuimg_particle_f");

  // Read side of particle_container<F, particle_f<T>, A> for one scale of a
  // device tracker.  Accessors copy the device state on demand.
  class particle_container_view
  {
  public:
    typedef particle_f<translation_transform> particle_type;
    typedef std::vector<particle_type> particle_vector;
    typedef std::vector<unsigned int> uint_vector;

    particle_container_view(std::function<cuimg_hip_tracker*()> t, std::function<hip_context*()> ctx, int scale)
      : t_(t), ctx_(ctx), scale_(scale) {}

    const particle_vector& dense_particles() const
    {
      cuimg_hip_tracker* t = t_();
      uint32_t n = 0;
      dense_.clear();
      if (!t) return dense_;
      ctx_()->check(cuimg_hip_tracker_get_particles(t, scale_, nullptr, 0, &n), "dense_particles");
      dense_.resize(n);
      if (n)
        ctx_()->check(cuimg_hip_tracker_get_particles(t, scale_, reinterpret_cast<cuimg_particle_f*>(dense_.data()),
                                                    n, &n), "dense_particles");
      return dense_;
    }

    int size() const
    {
      cuimg_hip_tracker* t = t_();
      uint32_t n = 0;
      if (t) ctx_()->check(cuimg_hip_tracker_nparticles(t, scale_, &n), "nparticles");
      return (int)n;
    }

    const uint_vector& matches() const
    {
      cuimg_hip_tracker* t = t_();
      uint32_t n = 0;
      matches_.clear();
      if (!t) return matches_;
      ctx_()->check(cuimg_hip_tracker_get_matches(t, scale_, nullptr, 0, &n), "matches");
      matches_.resize(n);
      if (n) ctx_()->check(cuimg_hip_tracker_get_matches(t, scale_, matches_.data(), n, &n), "matches");
      return matches_;
    }

    bool compact_has_run() const
    {
      cuimg_hip_tracker* t = t_();
      int v = 0;
      if (t) ctx_()->check(cuimg_hip_tracker_compact_has_run(t, scale_, &v), "compact_has_run");
      return v != 0;
    }

    // sparse_particles(): index image, 0xFFFFFFFF = empty.
    host_image2d<unsigned int> sparse_particles() const
    {
      cuimg_img v;
      ctx_()->check(cuimg_hip_tracker_sparse(t_(), scale_, &v), "sparse_particles");
      host_image2d<unsigned int> out(v.nrows, v.ncols);
      ctx_()->check(cuimg_hip_image_download(ctx_()->get(), &v, out.begin(), out.pitch()), "sparse_particles");
      return out;
    }

    // sync_attributes (tracking2/particle_container.hpp:511-545): keeps a
    // third-party attribute vector aligned with the dense vector through the
    // last compaction (matches: old index -> new index, -1 = died).
    template <typename T, typename D>
    void sync_attributes(T& v, typename T::value_type new_value, D die_fun) const
    {
      const unsigned nparts = (unsigned)size();
      if (compact_has_run())
      {
        const uint_vector& m = matches();
        T tmp(nparts, new_value);
        for (unsigned i = 0; i < m.size(); i++)
        {
          int ni = (int)m[i];
          if (ni >= 0)
          {
            if (i < v.size()) tmp[ni] = std::move(v[i]);
          }
          else if (i < v.size())
            die_fun(v[i]);
        }
        v.swap(tmp);
      }
      else
        v.resize(nparts, new_value);
    }

    template <typename T>
    void sync_attributes(T& v, typename T::value_type new_value = typename T::value_type()) const
    {
      sync_attributes(v, new_value, [](const typename T::value_type&) {});
    }

  private:
    std::function<cuimg_hip_tracker*()> t_;
    std::function<hip_context*()> ctx_;
    int scale_;
    mutable particle_vector dense_;
    mutable uint_vector matches_;
  };
}

#endif
