// Oracle: FP64 photometric LK loss and its exact gradient.
// TEST INFRASTRUCTURE ONLY (see oracle_core.hpp).
//
//   L = sum_p sum_w ( (G*I1)(x_p + d_p + w) - (G*I0)(x_p + w) )^2
//
// G*I is the separable clamp-to-edge Gaussian of gaussian_sep (same taps, same
// right-to-left order), samples are the bilinear `li` of the matcher (base pixel
// truncated toward zero, clamp-to-edge reads), w runs over a ws x ws window.
// The loss is not in the reference; it is the photometric term its matcher
// minimises (tracking2/pyrlk_tracker.hpp:283-317), written as a function.
#include "oracle_core.hpp"

namespace orc {

namespace {

struct bil {
  int r0, c0, r1, c1;  // clamped corner rows / cols
  double a0, a1;
};

inline bil corners(const image<double>& img, double qr, double qc) {
  int xr = (int)qr, xc = (int)qc;  // toward zero
  bil b;
  b.a0 = qr - xr;
  b.a1 = qc - xc;
  auto cr = [&](int r) { return r < 0 ? 0 : (r >= img.nr ? img.nr - 1 : r); };
  auto cc = [&](int c) { return c < 0 ? 0 : (c >= img.nc ? img.nc - 1 : c); };
  b.r0 = cr(xr);
  b.r1 = cr(xr + 1);
  b.c0 = cc(xc);
  b.c1 = cc(xc + 1);
  return b;
}

inline double sample(const image<double>& img, const bil& b) {
  return (1 - b.a0) * (1 - b.a1) * img(b.r0, b.c0) + b.a0 * (1 - b.a1) * img(b.r1, b.c0) +
         (1 - b.a0) * b.a1 * img(b.r0, b.c1) + b.a0 * b.a1 * img(b.r1, b.c1);
}

// Adjoint of the separable clamp-to-edge convolution: the transposed taps,
// with every clamped read turned into an accumulation into the edge pixel.
void blur_adjoint(const image<double>& bbar, const double* coef, int K, image<double>& ibar) {
  const int nr = bbar.nr, nc = bbar.nc;
  image<double> tbar(nr, nc, 0.0);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      double g = bbar(r, c);
      if (g == 0.0) continue;
      for (int j = -K; j <= K; j++) {
        int rr = r + j;
        rr = rr < 0 ? 0 : (rr >= nr ? nr - 1 : rr);
        tbar(rr, c) += coef[j + K] * g;
      }
    }
  ibar = image<double>(nr, nc, 0.0);
  for (int r = 0; r < nr; r++)
    for (int c = 0; c < nc; c++) {
      double g = tbar(r, c);
      if (g == 0.0) continue;
      for (int j = -K; j <= K; j++) {
        int cc = c + j;
        cc = cc < 0 ? 0 : (cc >= nc ? nc - 1 : cc);
        ibar(r, cc) += coef[j + K] * g;
      }
    }
}

}  // namespace

double lk_loss_grad_f64(const image<double>& I0, const image<double>& I1, const double* coef,
                        int K, const double* pos2, const double* d2, int np, int ws,
                        double* dL_dd2, image<double>* dL_dI1) {
  image<double> B0, B1;
  gaussian_sep_f64(I0, coef, K, B0);
  gaussian_sep_f64(I1, coef, K, B1);
  const int hws = ws / 2;
  image<double> S(I1.nr, I1.nc, 0.0);
  double L = 0.0;
  for (int p = 0; p < np; p++) {
    double gr = 0.0, gc = 0.0;
    for (int r = -hws; r <= hws; r++)
      for (int c = -hws; c <= hws; c++) {
        double q1r = pos2[2 * p] + d2[2 * p] + r, q1c = pos2[2 * p + 1] + d2[2 * p + 1] + c;
        double q0r = pos2[2 * p] + r, q0c = pos2[2 * p + 1] + c;
        bil b1 = corners(B1, q1r, q1c), b0 = corners(B0, q0r, q0c);
        double res = sample(B1, b1) - sample(B0, b0);
        L += res * res;
        double v00 = B1(b1.r0, b1.c0), v10 = B1(b1.r1, b1.c0), v01 = B1(b1.r0, b1.c1),
               v11 = B1(b1.r1, b1.c1);
        double dvr = (1 - b1.a1) * (v10 - v00) + b1.a1 * (v11 - v01);
        double dvc = (1 - b1.a0) * (v01 - v00) + b1.a0 * (v11 - v10);
        gr += 2 * res * dvr;
        gc += 2 * res * dvc;
        if (dL_dI1) {
          double t = 2 * res;
          S(b1.r0, b1.c0) += t * (1 - b1.a0) * (1 - b1.a1);
          S(b1.r1, b1.c0) += t * b1.a0 * (1 - b1.a1);
          S(b1.r0, b1.c1) += t * (1 - b1.a0) * b1.a1;
          S(b1.r1, b1.c1) += t * b1.a0 * b1.a1;
        }
      }
    if (dL_dd2) {
      dL_dd2[2 * p] = gr;
      dL_dd2[2 * p + 1] = gc;
    }
  }
  if (dL_dI1) blur_adjoint(S, coef, K, *dL_dI1);
  return L;
}

}  // namespace orc
