// Oracle: FAST-like saliency.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <cstdlib>

namespace orc {

// Radius-3 Bresenham ring, CPU table order (cuimg/neighb2d_data.h:184-189),
// offsets in (row, col).
static const int kRing[16][2] = {
    {-3, 0}, {-3, 1}, {-2, 2}, {-1, 3}, {0, 3},  {1, 3},  {2, 2},   {3, 1},
    {3, 0},  {3, -1}, {2, -2}, {1, -3}, {0, -3}, {-1, -3}, {-2, -2}, {-3, -1}};

// fast::compute_saliency (tracking2/fast_detector.hpp:26-157).
//  * early rejection on ring points {0,4,8,12,0,4} for n >= 9 (:65-97);
//  * longest run of consecutive ring points farther than th from the centre with
//    one sign, the last run continuing through index 0 (:99-147);
//  * score = max(sum of bright distances, sum of dark distances) over all 16
//    points beyond th, returned as gl8u, i.e. modulo 256 (:155, cuimg/gl.h:84);
//  * th is the detector's float threshold converted to int at the call (:28,291).
uint8_t fast_score(const image<uint8_t>& in, int r, int c, int n_, int th) {
  unsigned max_n = 0;
  unsigned n = 0;
  int status = 2;
  const int vp = in.clamped(r, c);
  unsigned sum_bright = 0;
  unsigned sum_dark = 0;

  if (n_ >= 9) {
    static const int to_test[6] = {0, 4, 8, 12, 0, 4};
    int equals = 0;
    for (int i = 0; i < 6; i++) {
      const int* o = kRing[to_test[i]];
      int vn = in.clamped(r + o[0], c + o[1]);
      int sign = vn > vp;
      unsigned char dist = (unsigned char)std::abs(vn - vp);
      if (dist > th) {
        if (sign == status) {
          n++;
        } else {
          if (n > max_n) max_n = n;
          status = sign;
          n = 1;
        }
      } else {
        if (n > max_n) max_n = n;
        status = 2;
        equals++;
        if (i < 4 && n_ >= 12 && equals >= 2) return 0;
      }
    }
    if (n > max_n) max_n = n;
    if (n_ >= 12 && max_n < 3) return 0;
    else if (n_ >= 8 && max_n < 2) return 0;
  }

  n = 0;
  status = 2;
  max_n = 0;
  for (int i = 0; i < 16; i++) {
    int vn = in.clamped(r + kRing[i][0], c + kRing[i][1]);
    int sign = vn > vp;
    unsigned char dist = (unsigned char)std::abs(vn - vp);
    if (dist > th) {
      if (sign == status) {
        n++;
      } else {
        if (n > max_n) max_n = n;
        status = sign;
        n = 1;
      }
      if (vn < vp) sum_dark += dist;
      else sum_bright += dist;
    } else {
      if (n > max_n) max_n = n;
      status = 2;
    }
  }

  if (n != 16 && status != 2) {
    for (int i = 0; i < 16; i++) {
      int vn = in.clamped(r + kRing[i][0], c + kRing[i][1]);
      int sign = vn > vp;
      unsigned char dist = (unsigned char)std::abs(vn - vp);
      if (dist <= th || sign != status) break;
      n++;
    }
  }

  if (n > max_n) max_n = n;
  if (max_n < (unsigned)n_) return 0;
  return (uint8_t)std::max(sum_bright, sum_dark);
}

// compute_saliency_kernel + fast_detector::update (tracking2/fast_detector.hpp:267-313):
// saliency <- 0, then the score wherever mask != 0, over the whole input domain.
void fast_saliency(const image<uint8_t>& in, const image<uint8_t>& mask, int n, int th,
                   image<uint8_t>& sal) {
  sal = image<uint8_t>(in.nr, in.nc, 0);
#pragma omp parallel for schedule(static, 2)
  for (int r = 0; r < in.nr; r++)
    for (int c = 0; c < in.nc; c++)
      if (mask(r, c) != 0) sal(r, c) = fast_score(in, r, c, n, th);
}

}  // namespace orc
