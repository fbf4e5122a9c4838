// Oracle: integer Gaussians, contrast, bc2s descriptor and the gradient-descent
// matcher.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

namespace orc {

// gaussian_blur_rows_sigma1 / gaussian_blur_cols_sigma1 (cuimg/cpu/gaussian_blur.h:9-47):
// out = (2 v + left + right) / 4 in int, stored as gl8u (values stay <= 255).
// The reference leaves the first and last column of the row pass and the first
// and last row of the column pass unwritten (those pixels keep earlier buffer
// contents); here they use the same formula with the missing neighbour
// replaced by the edge pixel, so both passes are clamp-to-edge.
void blur121_u8(const image<uint8_t>& in, image<uint8_t>& out) {
  const int nr = in.nr, nc = in.nc;
  image<uint8_t> t(nr, nc);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++)
      t(r, c) = (uint8_t)((2 * (int)in(r, c) + (int)in.clamped(r, c - 1) + (int)in.clamped(r, c + 1)) / 4);
  out = image<uint8_t>(nr, nc);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++)
      out(r, c) = (uint8_t)((2 * (int)t(r, c) + (int)t.clamped(r - 1, c) + (int)t.clamped(r + 1, c)) / 4);
}

// gaussian_kernel<cuda_gpu> (cuimg/gpu/gaussian_blur.h:31-53): tap i is
// gaussian_derivative<0>(sigma, i) * 100 truncated to int; the divisor is the
// sum of the taps.
std::vector<int> int_gauss_taps(float sigma, int hw) {
  std::vector<int> t(2 * hw + 1);
  for (int i = -hw; i <= hw; i++) t[i + hw] = (int)(gauss_coef_runtime(sigma, i) * 100.f);
  return t;
}

// gaussian_blur(src, dst, tmp, kernel) (cuimg/gpu/gaussian_blur.h:71-93): a
// column filter into a u8 temporary, then a row filter, each divided by the
// tap sum.  NPP's division rounding and anchor placement are not given by the
// reference; here each pass is centred, divides with truncation, and reads
// clamp-to-edge.
void gaussian_int_u8(const image<uint8_t>& in, float sigma, int hw, image<uint8_t>& out) {
  const std::vector<int> w = int_gauss_taps(sigma, hw);
  int sum = 0;
  for (int v : w) sum += v;
  const int nr = in.nr, nc = in.nc;
  image<uint8_t> t(nr, nc);
  out = image<uint8_t>(nr, nc);
  if (sum <= 0) return;
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      int a = 0;
      for (int k = -hw; k <= hw; k++) a += w[k + hw] * (int)in.clamped(r + k, c);
      t(r, c) = (uint8_t)std::min(255, a / sum);
    }
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      int a = 0;
      for (int k = -hw; k <= hw; k++) a += w[k + hw] * (int)t.clamped(r, c + k);
      out(r, c) = (uint8_t)std::min(255, a / sum);
    }
}

// contrast_kernel (tracking2/tracking_strategies.hpp:104-136), d = 2, on an
// input whose border is clamp-filled.
void contrast_u32(const image<uint8_t>& in, image<uint32_t>& out) {
  out = image<uint32_t>(in.nr, in.nc);
  auto v = [&](int r, int c) { return (int)in.clamped(r, c); };
  for (int r = 0; r < in.nr; r++)
    for (int c = 0; c < in.nc; c++)
      out(r, c) = (uint32_t)(std::abs(v(r, c + 2) - v(r, c - 2)) + std::abs(v(r - 2, c) - v(r + 2, c)) +
                             std::abs(v(r - 2, c - 2) - v(r + 2, c + 2)) +
                             std::abs(v(r - 2, c + 2) - v(r + 2, c - 2)));
}

// circle_r3_h[0, 2, ..., 14] (cuimg/neighb2d_data.h:184-189): the offsets of
// bc2s_feature's ctor (tracking2/bc2s_feature.hpp:200-216), (row, col).
static const int kRing8[8][2] = {{-3, 0}, {-2, 2}, {0, 3}, {2, 2}, {3, 0}, {2, -2}, {0, -3}, {-2, -2}};
// c8_h (cuimg/neighb2d_data.h:80-85) and c8_it_h (:162-180).
static const int kC8[8][2] = {{-1, 1}, {0, 1}, {1, 1}, {-1, 0}, {1, 0}, {-1, -1}, {0, -1}, {1, -1}};
static const int kC8It[9][2] = {{6, 3}, {0, 3}, {0, 5}, {2, 5}, {2, 7}, {4, 7}, {4, 1}, {6, 1}, {0, 0}};

// bc2s_internals::compute_feature (tracking2/bc2s_feature.hpp:119-143); s1 has
// a clamp-filled border of 3, s2 of 6, which the offsets stay inside.
feat16 bc2s_feature(const bc2s_images& F, int r, int c) {
  feat16 f;
  for (int i = 0; i < 8; i++) {
    f.d[i] = F.s1->clamped(r + kRing8[i][0], c + kRing8[i][1]);
    f.d[8 + i] = F.s2->clamped(r + 2 * kRing8[i][0], c + 2 * kRing8[i][1]);
  }
  return f;
}

// bc2s_internals::distance (tracking2/bc2s_feature.hpp:33-96): with every
// weight 255, (sum 255 |v - a|) / 255 is the plain sum of absolute
// differences; scale 1 adds the s1 ring to the s2 ring, any other scale uses
// the s2 ring only.
int bc2s_distance(const bc2s_images& F, const feat16& a, int r, int c, int scale) {
  int d = 0, d2 = 0;
  if (scale == 1)
    for (int i = 0; i < 8; i++)
      d += 255 * std::abs((int)F.s1->clamped(r + kRing8[i][0], c + kRing8[i][1]) - (int)a.d[i]);
  for (int i = 0; i < 8; i++)
    d2 += 255 * std::abs((int)F.s2->clamped(r + 2 * kRing8[i][0], c + 2 * kRing8[i][1]) - (int)a.d[8 + i]);
  return (d + d2) / 255;
}

// gradient_descent_match (tracking2/gradient_descent_matcher.hpp:9-74): at most
// 10 moves to the best of the 8 neighbours (after the first move only the 3 or
// 5 neighbours on the side of the last improving one, per c8_it_h), each move
// accepted when strictly better than the best distance so far.  The starting
// distance is the scale-1 distance whatever the scale.  A prediction outside
// the domain returns the prediction itself as the displacement, distance
// 999999.
descent_result gradient_descent_match(const bc2s_images& F, int pr0, int pc0, const feat16& f, int scale) {
  const int nr = F.s1->nr, nc = F.s1->nc;
  auto has = [&](int r, int c) { return r >= 0 && r < nr && c >= 0 && c < nc; };
  int pr = pr0, pc = pc0;
  int mr = pr, mc = pc;
  float md = (float)bc2s_distance(F, f, pr, pc, 1);
  int mi = 8;
  if (!has(pr, pc)) return descent_result{pr0, pc0, 999999.f};
  for (int search = 0; search < 10; search++) {
    int i = kC8It[mi][0];
    const int end = kC8It[mi][1];
    auto eval = [&](int k) {
      const int r = pr + kC8[k][0], c = pc + kC8[k][1];
      if (has(r, c)) {
        const float d = (float)bc2s_distance(F, f, r, c, scale);
        if (d < md) {
          mr = r;
          mc = c;
          mi = k;
          md = d;
        }
      }
    };
    eval(i);
    i = (i + 1) & 7;
    for (; i != end; i = (i + 1) & 7) eval(i);
    if ((pr == mr && pc == mc) || !has(mr, mc)) break;
    pr = mr;
    pc = mc;
  }
  return descent_result{mr - pr0, mc - pc0, md};
}

// two_step_gradient_descent_match (tracking2/gradient_descent_matcher.hpp:117-125):
// a descent on the s2 ring, then one on both rings from where it ended.
descent_result two_step_gradient_descent_match(const bc2s_images& F, int pr, int pc, const feat16& f) {
  const descent_result m1 = gradient_descent_match(F, pr, pc, f, 2);
  const descent_result m2 = gradient_descent_match(F, pr + m1.dr, pc + m1.dc, f, 1);
  return descent_result{m1.dr + m2.dr, m1.dc + m2.dc, m2.dist};
}

}  // namespace orc
