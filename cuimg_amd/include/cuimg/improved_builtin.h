// Minimal (row, col) vector types of the reference (cuimg/improved_builtin.h):
// x is the row, y the column; float -> int conversion truncates toward zero.
#ifndef CUIMG_IMPROVED_BUILTIN_H_
# define CUIMG_IMPROVED_BUILTIN_H_

namespace cuimg
{
  template <typename T>
  struct vec2
  {
    T x, y;
    vec2() : x(), y() {}
    vec2(T r, T c) : x(r), y(c) {}
    template <typename U>
    vec2(const vec2<U>& o) : x(static_cast<T>(o.x)), y(static_cast<T>(o.y)) {}
    T& r() { return x; }
    T& c() { return y; }
    const T& r() const { return x; }
    const T& c() const { return y; }
    T& operator[](int i) { return i ? y : x; }
    const T& operator[](int i) const { return i ? y : x; }
  };

  template <typename T> vec2<T> operator+(vec2<T> a, vec2<T> b) { return vec2<T>(a.x + b.x, a.y + b.y); }
  template <typename T> vec2<T> operator-(vec2<T> a, vec2<T> b) { return vec2<T>(a.x - b.x, a.y - b.y); }
  template <typename T> bool operator==(vec2<T> a, vec2<T> b) { return a.x == b.x && a.y == b.y; }
  template <typename T> bool operator!=(vec2<T> a, vec2<T> b) { return !(a == b); }

  typedef vec2<int> i_int2;
  typedef vec2<short> i_short2;
  typedef vec2<float> i_float2;
}

#endif
