// Domains and borders with the reference's semantics (cuimg/obox2d.h:29-78,
// cuimg/box2d.hpp:13-71, cuimg/border.h:12-58).
#ifndef CUIMG_OBOX2D_H_
# define CUIMG_OBOX2D_H_

# include <cuimg/improved_builtin.h>

namespace cuimg
{
  // Origin-anchored domain [0, nrows) x [0, ncols).
  class obox2d
  {
  public:
    obox2d() : nrows_(0), ncols_(0) {}
    obox2d(unsigned nrows, unsigned ncols) : nrows_(nrows), ncols_(ncols) {}
    int nrows() const { return nrows_; }
    int ncols() const { return ncols_; }
    int size() const { return nrows_ * ncols_; }
    bool has(const i_int2& p) const { return p.x >= 0 && p.x < nrows_ && p.y >= 0 && p.y < ncols_; }
  private:
    int nrows_, ncols_;
  };

  inline bool operator==(const obox2d& a, const obox2d& b) { return a.nrows() == b.nrows() && a.ncols() == b.ncols(); }
  inline bool operator!=(const obox2d& a, const obox2d& b) { return !(a == b); }
  // obox2d / x truncates (the float quotient converted to unsigned)
  inline obox2d operator/(const obox2d& a, float x) { return obox2d(unsigned(a.nrows() / x), unsigned(a.ncols() / x)); }

  // Inclusive box [p1, p2].
  class box2d
  {
  public:
    box2d() {}
    box2d(i_int2 a, i_int2 b) : a_(a), b_(b) {}
    box2d(const obox2d& d) : a_(0, 0), b_(d.nrows() - 1, d.ncols() - 1) {}
    const i_int2& p1() const { return a_; }
    const i_int2& p2() const { return b_; }
    bool has(const i_int2& p) const { return p.x >= a_.x && p.x <= b_.x && p.y >= a_.y && p.y <= b_.y; }
  private:
    i_int2 a_, b_;
  };

  class border
  {
  public:
    border(int t = 0) : t_(t) {}
    int thickness() const { return t_; }
  private:
    int t_;
  };

  inline box2d operator-(const obox2d& d, const border& b)
  {
    int t = b.thickness();
    return box2d(i_int2(t, t), i_int2(d.nrows() - t - 1, d.ncols() - t - 1));
  }
}

#endif
