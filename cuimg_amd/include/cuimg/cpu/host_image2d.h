// host_image2d<V>: host image with the reference's layout and handle semantics
// (cuimg/cpu/host_image2d.h:29-118, .hpp:78-112): (nrows + 2*border) rows of
// `pitch` bytes, pitch = (ncols + 2*border)*sizeof(V) rounded up to 4, begin()
// at pixel (0,0); copies are shallow (shared ownership), copy() is deep.
#ifndef CUIMG_HOST_IMAGE2D_H_
# define CUIMG_HOST_IMAGE2D_H_

# include <cstring>
# include <memory>

# include <cuimg/obox2d.h>

namespace cuimg
{
  template <typename V>
  class host_image2d
  {
  public:
    typedef V value_type;
    typedef obox2d domain_type;

    host_image2d() : pitch_(0), border_(0), begin_(nullptr) {}
    host_image2d(unsigned nrows, unsigned ncols, unsigned border = 0) { allocate(obox2d(nrows, ncols), border); }
    host_image2d(const obox2d& d, unsigned border = 0) { allocate(d, border); }

    const obox2d& domain() const { return domain_; }
    int nrows() const { return domain_.nrows(); }
    int ncols() const { return domain_.ncols(); }
    int border() const { return border_; }
    size_t pitch() const { return pitch_; }
    bool has(const i_int2& p) const { return domain_.has(p); }

    V* begin() const { return begin_; }
    V* row(int r) { return reinterpret_cast<V*>(reinterpret_cast<char*>(begin_) + (ptrdiff_t)r * pitch_); }
    const V* row(int r) const
    {
      return reinterpret_cast<const V*>(reinterpret_cast<const char*>(begin_) + (ptrdiff_t)r * pitch_);
    }
    V& operator()(int r, int c) { return row(r)[c]; }
    const V& operator()(int r, int c) const { return row(r)[c]; }
    V& operator()(const i_int2& p) { return row(p.x)[p.y]; }
    const V& operator()(const i_int2& p) const { return row(p.x)[p.y]; }

  private:
    void allocate(const obox2d& d, unsigned border)
    {
      domain_ = d;
      border_ = border;
      pitch_ = (d.ncols() + 2 * border) * sizeof(V);
      if (pitch_ % 4) pitch_ += 4 - (pitch_ & 3);
      size_t bytes = (size_t)(d.nrows() + 2 * border) * pitch_ + 64;
      data_ = std::shared_ptr<char>(new char[bytes](), std::default_delete<char[]>());
      begin_ = reinterpret_cast<V*>(data_.get() + border * pitch_) + border;
    }

    obox2d domain_;
    size_t pitch_;
    int border_;
    std::shared_ptr<char> data_;
    V* begin_;
  };

  // Deep copy of the domain (cuimg/copy.h:31-40, without the equal-border
  // shortcut that skips the last rows).
  template <typename V>
  void copy(const host_image2d<V>& in, host_image2d<V>& out)
  {
    for (int r = 0; r < in.nrows(); r++) std::memcpy(out.row(r), in.row(r), in.ncols() * sizeof(V));
  }
}

#endif
