// Particle container of one tracker scale, read from the device
// (tracking2/particle_container.h:13-225, .hpp:22-578).  The dense AoS layout
// particle_f<translation_transform> is the reference's: 20 bytes,
// {speed(row,col), pos(row,col), unsigned short age}; the bc2s strategy's
// particle (integer position) is converted from it.
#ifndef CUIMG_TRACKING2_PARTICLE_CONTAINER_H_
# define CUIMG_TRACKING2_PARTICLE_CONTAINER_H_

# include <cstring>
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

  // particle (tracking2/particle_container.h:13-45): the bc2s strategy's
  // particle, integer position.  The match times are not used by the
  // strategy and read as 0.
  struct particle
  {
    typedef translation_transform transform;
    typedef i_short2 coords_type;
    particle() : speed(0.f, 0.f), age(0), fault(0), prev_match_time(0), next_match_time(0) {}
    translation_transform speed;
    i_short2 pos;
    unsigned short age;
    unsigned short fault;
    int prev_match_time;
    int next_match_time;
  };

  // bc2s (tracking2/bc2s_feature.h:16-43): 16 descriptor bytes and their
  // weights (255 for every descriptor the strategy builds).
  struct bc2s
  {
    bc2s() { for (int i = 0; i < 16; i++) distances[i] = 0, weights[i] = 255; }
    const unsigned char& operator[](unsigned i) const { return distances[i]; }
    unsigned char& operator[](unsigned i) { return distances[i]; }
    unsigned char distances[16];
    unsigned char weights[16];
  };

  namespace internal
  {
    inline void from_device(const cuimg_particle_f& d, particle_f<translation_transform>& p)
    {
      static_assert(sizeof(p) == sizeof(d), "same layout");
      std::memcpy(static_cast<void*>(&p), &d, sizeof d);
    }
    inline void from_device(const cuimg_particle_f& d, particle& p)
    {
      p.speed = translation_transform(d.speed_r, d.speed_c);
      p.pos = i_short2((short)d.pos_r, (short)d.pos_c);
      p.age = d.age;
      p.fault = 0;
      p.prev_match_time = p.next_match_time = 0;
    }
  }

  // Read side of particle_container<F, P, A> for one scale of a device
  // tracker.  Accessors copy the device state on demand.
  template <typename P>
  class particle_container_view_t
  {
  public:
    typedef P particle_type;
    typedef std::vector<particle_type> particle_vector;
    typedef std::vector<unsigned int> uint_vector;

    particle_container_view_t(std::function<cuimg_hip_tracker*()> t, std::function<hip_context*()> ctx, int scale)
      : t_(t), ctx_(ctx), scale_(scale) {}

    const particle_vector& dense_particles() const
    {
      cuimg_hip_tracker* t = t_();
      uint32_t n = 0;
      dense_.clear();
      if (!t) return dense_;
      ctx_()->check(cuimg_hip_tracker_get_particles(t, scale_, nullptr, 0, &n), "dense_particles");
      raw_.resize(n);
      if (n) ctx_()->check(cuimg_hip_tracker_get_particles(t, scale_, raw_.data(), n, &n), "dense_particles");
      dense_.resize(n);
      for (uint32_t i = 0; i < n; i++) internal::from_device(raw_[i], dense_[i]);
      return dense_;
    }

    // particle_container::features() of the bc2s strategy.
    const std::vector<bc2s>& features() const
    {
      cuimg_hip_tracker* t = t_();
      uint32_t n = 0;
      feats_.clear();
      if (!t) return feats_;
      ctx_()->check(cuimg_hip_tracker_get_features(t, scale_, nullptr, 0, &n), "features");
      std::vector<uint8_t> bytes(16 * (size_t)n);
      if (n) ctx_()->check(cuimg_hip_tracker_get_features(t, scale_, bytes.data(), n, &n), "features");
      feats_.resize(n);
      for (uint32_t i = 0; i < n; i++) std::memcpy(feats_[i].distances, &bytes[16 * (size_t)i], 16);
      return feats_;
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
    mutable std::vector<cuimg_particle_f> raw_;
    mutable particle_vector dense_;
    mutable uint_vector matches_;
    mutable std::vector<bc2s> feats_;
  };

  typedef particle_container_view_t<particle_f<translation_transform> > particle_container_view;
}

#endif
