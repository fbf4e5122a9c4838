// gl8u: 8-bit gray level (cuimg/gl.h:14-95); arithmetic converts to unsigned char.
#ifndef CUIMG_GL_H_
# define CUIMG_GL_H_

namespace cuimg
{
  struct gl8u
  {
    unsigned char x;
    gl8u() : x(0) {}
    template <typename T> gl8u(const T& v) : x(static_cast<unsigned char>(v)) {}
    operator unsigned char() const { return x; }
  };
  static_assert(sizeof(gl8u) == 1, "gl8u is one byte");
}

#endif
