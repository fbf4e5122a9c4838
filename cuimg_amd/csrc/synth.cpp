// Seeded synthetic texture video for the benchmarks (host C++, std::thread).
//
// Not part of the reference (its sample reads a camera / video file through
// OpenCV); this generator replaces that input with a deterministic workload:
//   frame 0 : white noise N(0,1) -> normalised FP64 Gaussian, sigma 2, 13 taps,
//             clamp-to-edge -> affinely rescaled to [0,255] -> 20 axis-aligned
//             rectangles of uniform intensity painted in order -> rounded
//             (half to even) and clamped to uint8;
//   frame k : frame k-1 resampled (bilinear, clamp-to-edge, FP64) through a
//             random motion about the image centre, plus N(0, 2^2) noise,
//             rounded half to even, clamped to uint8.
// Motion 0: translation U[-3,3] px per axis.  Motion 1: rotation U[-2,2] deg,
// scale U[0.98,1.02], translation U[-8,8] px per axis.
// Random numbers are counter based (SplitMix64 finaliser of seed, stream and
// index), so every pixel is independent of the thread that computes it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/cuimg_hip.h"

namespace {

// Rows [0, n) split over the host threads.
template <typename F>
void parallel_rows(int n, F&& f) {
  int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 32u);
  nt = std::min(nt, std::max(1, n / 8));
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (int r = t; r < n; r += nt) f(r);
    });
  for (auto& x : th) x.join();
}

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t index) {
  return mix64(mix64(seed ^ mix64(stream)) + index);
}

inline double u01(uint64_t seed, uint64_t stream, uint64_t index) {
  return (double)(rnd(seed, stream, index) >> 11) * (1.0 / 9007199254740992.0);
}

inline double normal(uint64_t seed, uint64_t stream, uint64_t index) {
  double u1 = 1.0 - u01(seed, stream, 2 * index);  // (0, 1]
  double u2 = u01(seed, stream, 2 * index + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

inline uint8_t quantise(double v) {
  double r = std::nearbyint(v);
  return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
}

void texture(uint64_t seed, int nr, int nc, std::vector<double>& t) {
  const size_t n = (size_t)nr * nc;
  std::vector<double> noise(n), tmp(n);
  parallel_rows(nr, [&](int r) {
    for (int c = 0; c < nc; c++) noise[(size_t)r * nc + c] = normal(seed, 1, (uint64_t)r * nc + c);
  });
  const int K = 6;
  double g[2 * K + 1], sum = 0;
  for (int x = -K; x <= K; x++) sum += g[x + K] = std::exp(-x * x / 8.0);
  for (double& v : g) v /= sum;
  parallel_rows(nr, [&](int r) {
    for (int c = 0; c < nc; c++) {
      double s = 0;
      for (int j = -K; j <= K; j++) s += g[j + K] * noise[(size_t)r * nc + std::min(std::max(c + j, 0), nc - 1)];
      tmp[(size_t)r * nc + c] = s;
    }
  });
  t.assign(n, 0.0);
  parallel_rows(nr, [&](int r) {
    for (int c = 0; c < nc; c++) {
      double s = 0;
      for (int j = -K; j <= K; j++) s += g[j + K] * tmp[(size_t)std::min(std::max(r + j, 0), nr - 1) * nc + c];
      t[(size_t)r * nc + c] = s;
    }
  });
  double lo = t[0], hi = t[0];
  for (double v : t) {
    lo = std::min(lo, v);
    hi = std::max(hi, v);
  }
  const double scale = hi > lo ? 255.0 / (hi - lo) : 0.0;
  for (double& v : t) v = (v - lo) * scale;
  for (int k = 0; k < 20; k++) {
    int ra = (int)(u01(seed, 2, 5 * k + 0) * nr), rb = (int)(u01(seed, 2, 5 * k + 1) * nr);
    int ca = (int)(u01(seed, 2, 5 * k + 2) * nc), cb = (int)(u01(seed, 2, 5 * k + 3) * nc);
    double v = u01(seed, 2, 5 * k + 4) * 255.0;
    for (int r = std::min(ra, rb); r <= std::max(ra, rb); r++)
      for (int c = std::min(ca, cb); c <= std::max(ca, cb); c++) t[(size_t)r * nc + c] = v;
  }
}

}  // namespace

extern "C" int cuimg_synth_video(uint64_t seed, int nr, int nc, int nframes, int motion, uint8_t* out) {
  if (nr <= 0 || nc <= 0 || nframes <= 0 || !out || (motion != 0 && motion != 1)) return CUIMG_E_INVALID;
  const size_t n = (size_t)nr * nc;
  std::vector<double> t;
  texture(seed, nr, nc, t);
  for (size_t i = 0; i < n; i++) out[i] = quantise(t[i]);
  const double cr = (nr - 1) / 2.0, cc = (nc - 1) / 2.0;
  for (int k = 1; k < nframes; k++) {
    const uint8_t* prev = out + (size_t)(k - 1) * n;
    uint8_t* cur = out + (size_t)k * n;
    double th = 0, s = 1, tr, tc;
    if (motion == 0) {
      tr = -3 + 6 * u01(seed, 3, 8 * k + 0);
      tc = -3 + 6 * u01(seed, 3, 8 * k + 1);
    } else {
      th = (-2 + 4 * u01(seed, 3, 8 * k + 2)) * M_PI / 180.0;
      s = 0.98 + 0.04 * u01(seed, 3, 8 * k + 3);
      tr = -8 + 16 * u01(seed, 3, 8 * k + 0);
      tc = -8 + 16 * u01(seed, 3, 8 * k + 1);
    }
    // out(p) = prev(M^-1 (p - c - t) + c), M = s R(th)
    const double ci = std::cos(th) / s, si = std::sin(th) / s;
    parallel_rows(nr, [&](int r) {
      for (int c = 0; c < nc; c++) {
        double y = r - cr - tr, x = c - cc - tc;
        double qr = ci * y + si * x + cr, qc = -si * y + ci * x + cc;
        int r0 = (int)std::floor(qr), c0 = (int)std::floor(qc);
        double ar = qr - r0, ac = qc - c0;
        auto at = [&](int rr, int c2) {
          rr = std::min(std::max(rr, 0), nr - 1);
          c2 = std::min(std::max(c2, 0), nc - 1);
          return (double)prev[(size_t)rr * nc + c2];
        };
        double v = (1 - ar) * ((1 - ac) * at(r0, c0) + ac * at(r0, c0 + 1)) +
                   ar * ((1 - ac) * at(r0 + 1, c0) + ac * at(r0 + 1, c0 + 1));
        v += 2.0 * normal(seed, 4, (uint64_t)k * n + (uint64_t)r * nc + c);
        cur[(size_t)r * nc + c] = quantise(v);
      }
    });
  }
  return 0;
}
