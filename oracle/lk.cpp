// Oracle: pyramid Lucas-Kanade matcher.  TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
#include "oracle_core.hpp"

#include <algorithm>
#include <cmath>

namespace orc {

// Bilinear sample `li` (tracking2/pyrlk_tracker.hpp:130-146): base pixel is the
// coordinate truncated toward zero, weights are the (possibly negative)
// remainders, terms summed left to right.  fused: the last three terms are
// added with fused multiply-adds (the GPU's arithmetic); otherwise every
// product is rounded first.
static inline float blend4(float w00, float w10, float w01, float w11, float v00, float v10, float v01, float v11,
                           bool fused) {
  float s = w00 * v00;
  if (fused) {
    s = std::fmaf(w10, v10, s);
    s = std::fmaf(w01, v01, s);
    return std::fmaf(w11, v11, s);
  }
  s = s + w10 * v10;
  s = s + w01 * v01;
  return s + w11 * v11;
}

static inline float li(const image<float>& img, f2 p, bool fused) {
  int xr = tz(p.r), xc = tz(p.c);
  float a0 = p.r - (float)xr;
  float a1 = p.c - (float)xc;
  float v00 = img.clamped(xr, xc), v10 = img.clamped(xr + 1, xc);
  float v01 = img.clamped(xr, xc + 1), v11 = img.clamped(xr + 1, xc + 1);
  float w00 = (1 - a0) * (1 - a1), w10 = a0 * (1 - a1), w01 = (1 - a0) * a1, w11 = a0 * a1;
  return blend4(w00, w10, w01, w11, v00, v10, v01, v11, fused);
}

static inline f2 li2(const image<f2>& img, f2 p, bool fused) {
  int xr = tz(p.r), xc = tz(p.c);
  float a0 = p.r - (float)xr;
  float a1 = p.c - (float)xc;
  f2 v00 = img.clamped(xr, xc), v10 = img.clamped(xr + 1, xc);
  f2 v01 = img.clamped(xr, xc + 1), v11 = img.clamped(xr + 1, xc + 1);
  float w00 = (1 - a0) * (1 - a1), w10 = a0 * (1 - a1), w01 = (1 - a0) * a1, w11 = a0 * a1;
  return f2{blend4(w00, w10, w01, w11, v00.r, v10.r, v01.r, v11.r, fused),
            blend4(w00, w10, w01, w11, v00.c, v10.c, v01.c, v11.c, fused)};
}

// Sobel/8 gradient of the previous frame (tracking2/pyrlk_tracker.hpp:909-938):
// zero on the 1-pixel outer ring, [0] = d/drow, [1] = d/dcol.  For integer
// sources the sum is an int and the division exact, so FP32 order is
// irrelevant; for float sources the same terms are summed left to right in FP32.
void sobel_gradient(const image<float>& P, bool integer_source, image<f2>& g) {
  g = image<f2>(P.nr, P.nc, f2{0.f, 0.f});
#pragma omp parallel for schedule(static, 2)
  for (int r = 1; r < P.nr - 1; r++)
    for (int c = 1; c < P.nc - 1; c++) {
      if (integer_source) {
        int gr = 1 * (int)P(r + 1, c - 1) + 2 * (int)P(r + 1, c) + 1 * (int)P(r + 1, c + 1) -
                 1 * (int)P(r - 1, c - 1) - 2 * (int)P(r - 1, c) - 1 * (int)P(r - 1, c + 1);
        int gc = 1 * (int)P(r - 1, c + 1) + 2 * (int)P(r, c + 1) + 1 * (int)P(r + 1, c + 1) -
                 1 * (int)P(r - 1, c - 1) - 2 * (int)P(r, c - 1) - 1 * (int)P(r + 1, c - 1);
        g(r, c) = f2{(float)(gr / 8.), (float)(gc / 8.)};
      } else {
        float gr = P(r + 1, c - 1) + 2 * P(r + 1, c);
        gr = gr + P(r + 1, c + 1);
        gr = gr - P(r - 1, c - 1);
        gr = gr - 2 * P(r - 1, c);
        gr = gr - P(r - 1, c + 1);
        float gc = P(r - 1, c + 1) + 2 * P(r, c + 1);
        gc = gc + P(r + 1, c + 1);
        gc = gc - P(r - 1, c - 1);
        gc = gc - 2 * P(r, c - 1);
        gc = gc - P(r + 1, c - 1);
        g(r, c) = f2{gr / 8.f, gc / 8.f};
      }
    }
}

// Window sums in a defined order.  The reference adds the terms in its loop
// order (raster, :231-250, :287-299, :320-333) with whatever contraction its
// compiler chose, so its rounding is not reproducible; the oracle offers two
// definitions:
//  * lanes == 0: raster order, each product rounded, then added;
//  * lanes == 16 or 64: window pixel i (row r = i / ws, column c = i % ws)
//    goes to lane 8 (r % 2) + c for 16 lanes (windows up to 7x7: a lane holds
//    one column at every other row) and to lane i % 64 for 64 lanes; each lane
//    accumulates its pixels in ascending order, a product term as
//    fma(a, b, acc); lane sums are then combined pairwise within rows of 16
//    lanes ((x0+x1)+(x2+x3)) + ((x4+x5)+(x6+x7)) and so on, and for 64 lanes
//    the four row sums as (r3 + r2) + (r1 + r0).
struct window_sum {
  int lanes, ws;
  float acc[64];
  float serial;
  window_sum(int l, int w) : lanes(l), ws(w), serial(0.f) { std::fill(acc, acc + 64, 0.f); }
  int lane(int i) const { return lanes == 16 ? ((i / ws) & 1) * 8 + i % ws : i % lanes; }
  void fma(int i, float a, float b) {
    if (lanes) acc[lane(i)] = std::fmaf(a, b, acc[lane(i)]);
    else serial = serial + a * b;
  }
  void add(int i, float x) {
    if (lanes) acc[lane(i)] = acc[lane(i)] + x;
    else serial = serial + x;
  }
  static float row16(const float* x) {
    const float p0 = x[0] + x[1], p1 = x[2] + x[3], p2 = x[4] + x[5], p3 = x[6] + x[7];
    const float p4 = x[8] + x[9], p5 = x[10] + x[11], p6 = x[12] + x[13], p7 = x[14] + x[15];
    return ((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7));
  }
  float total() const {
    if (!lanes) return serial;
    if (lanes == 16) return row16(acc);
    const float r0 = row16(acc), r1 = row16(acc + 16), r2 = row16(acc + 32), r3 = row16(acc + 48);
    return (r3 + r2) + (r1 + r0);
  }
};

// iterative_lk_match_5x5::match (tracking2/pyrlk_tracker.hpp:165-335),
// parameterised by window, iteration count, selection size and template type.
// Defined behaviour where the reference is undefined:
//  * window pixels outside the gradient image are excluded from selection and
//    from G (their gs/as are uninitialised in the reference, :188-206);
//  * selection of the top_n strongest |g| breaks ties by raster index
//    (std::nth_element's order is implementation-defined, :214);
//  * the eigenvalue test uses the closed form of a symmetric 2x2 matrix in
//    FP32 (Eigen's general solver, :252-258);
//  * G^-1 = adj(G)/det(G) with det = G00*G11 - G10*G01 (Eigen's 2x2 inverse).
// Kept quirks: uint8 truncation of template samples for u8 images (:190,200),
// the error's index (r+hws)*hws+(c+hws) and non-reset counter (:320-333),
// fail if |prediction - v| > 4*ws or v leaves domain - border(3) (:312-316).
match_result lk_match(const lk_params& P, f2 p, f2 pred_tr, const lk_level& L) {
  const match_result fail{{-1.f, -1.f}, 1000.f};
  const int ws = P.ws, hws = ws / 2, n = ws * ws;
  const image<float>& A = *L.A;
  const image<float>& B = *L.B;
  const image<f2>& Ag = *L.G;
  // the GPU's arithmetic (fused bilinear blends) goes with its window-sum order
  const bool fused = P.group_lanes != 0;

  f2 prediction{p.r + pred_tr.r, p.c + pred_tr.c};
  f2 v = prediction;
  float nk0 = 1.f, nk1 = 1.f;

  std::vector<f2> gs(n, f2{0.f, 0.f});
  std::vector<float> as(n, 0.f);
  std::vector<char> valid(n, 0), mask(n, 0);
  for (int i = 0, r = -hws; r <= hws; r++)
    for (int c = -hws; c <= hws; c++, i++) {
      f2 q{p.r + (float)r, p.c + (float)c};
      if (Ag.has(tz(q.r), tz(q.c))) {
        valid[i] = 1;
        gs[i] = li2(Ag, q, fused);
        float a = li(A, q, fused);
        // float -> uint8 through int (the x86 conversion; values outside
        // [0,255] only arise from the negative weights of truncated negative
        // coordinates and then wrap modulo 256).
        as[i] = P.template_u8 ? (float)(uint8_t)(int)a : a;
      }
    }

  if (P.top_n >= n) {
    for (int i = 0; i < n; i++) mask[i] = valid[i];
  } else {
    std::vector<int> idx;
    for (int i = 0; i < n; i++)
      if (valid[i]) idx.push_back(i);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
      return norml2(gs[a].r, gs[a].c) > norml2(gs[b].r, gs[b].c);
    });
    for (int k = 0; k < (int)idx.size() && k < P.top_n; k++) mask[idx[k]] = 1;
  }

  window_sum s00(P.group_lanes, ws), s01(P.group_lanes, ws), s11(P.group_lanes, ws);
  int cpt = 0;
  for (int i = 0; i < n; i++)
    if (mask[i]) {
      float gx = gs[i].r, gy = gs[i].c;
      s00.fma(i, gx, gx);
      s01.fma(i, gx, gy);
      s11.fma(i, gy, gy);
      cpt++;
    }
  // G is symmetric: G10 = G01 term by term in either order
  const float G00 = s00.total(), G01 = s01.total(), G10 = G01, G11 = s11.total();

  {
    float h = 0.5f * (G00 + G11);
    float q = 0.5f * (G00 - G11);
    float rr = std::sqrt(q * q + G01 * G10);
    float l1 = h + rr, l2 = h - rr;
    float min_ev = std::min(std::fabs(l1), std::fabs(l2));
    if (min_ev < P.min_eig) return fail;
  }

  float det = G00 * G11 - G10 * G01;
  float invdet = 1.f / det;
  float I00 = G11 * invdet, I01 = -G01 * invdet, I10 = -G10 * invdet, I11 = G00 * invdet;

  const box dom = inner(B.nr, B.nc, 3);
  for (int k = 0; k < P.max_iters && std::sqrt(nk0 * nk0 + nk1 * nk1) >= P.eps; k++) {
    window_sum sb0(P.group_lanes, ws), sb1(P.group_lanes, ws);
    cpt = 0;
    for (int i = 0, r = -hws; r <= hws; r++)
      for (int c = -hws; c <= hws; c++, i++)
        if (mask[i]) {
          f2 q{v.r + (float)r, v.c + (float)c};
          float dt = as[i] - li(B, q, fused);
          sb0.fma(i, gs[i].r, dt);
          sb1.fma(i, gs[i].c, dt);
          cpt++;
        }
    const float b0 = sb0.total(), b1 = sb1.total();
    nk0 = I00 * b0 + I01 * b1;
    nk1 = I10 * b0 + I11 * b1;
    v.r += nk0;
    v.c += nk1;
    if (norml2(prediction.r - v.r, prediction.c - v.c) > ws * 4) return fail;
    if (!dom.has(tz(v.r), tz(v.c))) return fail;
  }

  window_sum se(P.group_lanes, ws);
  for (int w = 0, r = -hws; r <= hws; r++)
    for (int c = -hws; c <= hws; c++, w++) {
      f2 q{v.r + (float)r, v.c + (float)c};
      int i = (r + hws) * hws + (c + hws);
      se.add(w, std::fabs(as[i] - li(B, q, fused)));
      cpt++;
    }
  return match_result{{v.r - p.r, v.c - p.c}, se.total() / cpt};
}

}  // namespace orc
