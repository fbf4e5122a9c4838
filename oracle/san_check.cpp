// Sanitizer driver of the CPU oracle (test infrastructure only): calls every
// entry point of oracle.h on small seeded inputs, so that a build with
// -fsanitize=address,undefined (or thread) checks the oracle's memory use and
// arithmetic for undefined behaviour.  A few of the invariants the Python
// tests hold are checked on the way (the frame step through update_input +
// step equals run; compaction keeps the live particles; the merge keeps the
// sparse image consistent), so a silent corruption also fails.
//
//   san_check [frames]     exit status 0 = clean; sanitizer reports abort
// (profiles 0 and 1: the pyramid-LK strategy; 2: the bc2s strategy)
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

uint64_t lcg(uint64_t& s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return s >> 33;
}

// Smooth texture (sum of a few separable cosines plus noise) moved by an
// integer translation per frame; values 0..255.
std::vector<uint8_t> texture_video(uint64_t seed, int nr, int nc, int nframes) {
  std::vector<uint8_t> v((size_t)nframes * nr * nc);
  std::vector<float> fr(6), fc(6), ph(6);
  for (int k = 0; k < 6; k++) {
    fr[k] = 0.05f + 0.3f * (float)(lcg(seed) % 1000) / 1000.f;
    fc[k] = 0.05f + 0.3f * (float)(lcg(seed) % 1000) / 1000.f;
    ph[k] = (float)(lcg(seed) % 6283) / 1000.f;
  }
  for (int f = 0; f < nframes; f++)
    for (int r = 0; r < nr; r++)
      for (int c = 0; c < nc; c++) {
        const float y = (float)(r + f), x = (float)(c + 2 * f);
        float a = 128.f;
        for (int k = 0; k < 6; k++) a += 18.f * std::cos(fr[k] * y + ph[k]) * std::cos(fc[k] * x + 0.5f * ph[k]);
        a += (float)(lcg(seed) % 9) - 4.f;
        v[((size_t)f * nr + r) * nc + c] = (uint8_t)(a < 0.f ? 0.f : (a > 255.f ? 255.f : a));
      }
  return v;
}

int failures = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    failures++;
  }
}

orc_tracker_params params(int profile, int nscales, float th, int lanes) {
  orc_tracker_params p{};
  p.nscales = nscales;
  p.profile = profile;
  p.lk_ws = profile == 0 ? 11 : 7;
  p.lk_max_iters = profile == 0 ? 30 : 5;
  p.lk_top_n = profile == 0 ? 25 : 49;
  p.lk_eps = 0.1f;
  p.fast_n = 9;
  p.fast_th = th;
  p.box_size = 3;
  p.detector_frequency = 1;
  p.filtering_frequency = 1;
  p.k = 300;
  p.with_merge = 1;
  p.blur_sigma = 1;
  p.blur_K = 3;
  p.sum_lanes = lanes;
  return p;
}

void images(const std::vector<uint8_t>& v, int nr, int nc) {
  const size_t N = (size_t)nr * nc;
  std::vector<float> f(v.begin(), v.begin() + N), f4(4 * N), out(N), out4(4 * N);
  for (size_t i = 0; i < 4 * N; i++) f4[i] = (float)v[i % N] * 0.5f + (float)(i % 4);
  std::vector<double> d(f.begin(), f.end()), outd(N);
  for (int sigma = 1; sigma <= 4; sigma *= 2) {
    const int K = 3 * sigma;
    std::vector<float> coef(2 * K + 1), coef_rt(2 * K + 1);
    orc_gauss_coefs(sigma, K, 1, coef.data());
    orc_gauss_coefs(sigma, K, 0, coef_rt.data());
    check(orc_gauss_coef_meta(sigma, 0) > 0.f && orc_gauss_coef_runtime((float)sigma, 1) > 0.f, "coefficients");
    std::vector<double> cd(coef.begin(), coef.end());
    for (int fused = 0; fused < 2; fused++) {
      orc_gaussian_sep_f32(f.data(), nr, nc, 1, coef.data(), K, fused, out.data());
      orc_gaussian_sep_f32(f4.data(), nr, nc, 4, coef.data(), K, fused, out4.data());
      orc_gaussian_sep_u8(v.data(), nr, nc, coef.data(), K, fused, out.data());
    }
    orc_gaussian_sep_f64(d.data(), nr, nc, 1, cd.data(), K, outd.data());
    double diff = 0.0;
    for (size_t i = 0; i < N; i++) diff = std::fmax(diff, std::fabs(outd[i] - out[i]));
    check(diff < 1e-2, "u8 and f64 Gaussians agree");
  }
  std::vector<uint8_t> h((size_t)(nr / 2) * (nc / 2));
  std::vector<float> hf(h.size());
  orc_pyr_down_u8(v.data(), nr, nc, h.data());
  orc_pyr_down_f32(f.data(), nr, nc, hf.data());
  std::vector<uint8_t> sal(N), mask(N, 255);
  orc_fast_saliency(v.data(), nr, nc, mask.data(), 9, 12, sal.data());
  check(orc_fast_score(v.data(), nr, nc, nr / 2, nc / 2, 9, 12) == sal[(size_t)(nr / 2) * nc + nc / 2],
        "fast score = saliency");
  std::vector<int32_t> sparse(N, -1), pts(2 * N);
  const int npts = orc_fast_select(sal.data(), sparse.data(), nr, nc, 3, 12.f, pts.data(), (int)N);
  check(npts > 0, "fast select finds points");
  std::vector<float> g(2 * N);
  orc_sobel_gradient(f.data(), nr, nc, 1, g.data());
  orc_sobel_gradient(f.data(), nr, nc, 0, g.data());
  // LK on one level: raster and lane-group sums, both templates
  for (int lanes : {0, 16, 64})
    for (int u8 = 0; u8 < 2; u8++) {
      float tr[2], dist;
      orc_lk_match(f.data(), f.data(), g.data(), nr, nc, u8 ? 11 : 7, u8 ? 30 : 5, 0.1f, u8 ? 25 : 49, u8, lanes,
                   nr / 2.f + 0.25f, nc / 3.f, nr / 2.f + 0.75f, nc / 3.f - 0.5f, tr, &dist);
      check(std::isfinite(tr[0]) && std::isfinite(tr[1]) && std::isfinite(dist), "lk match finite");
      // near the border: the clamped reads
      orc_lk_match(f.data(), f.data(), g.data(), nr, nc, 7, 5, 0.1f, 49, u8, lanes, 3.f, 3.f, 3.9f, nc - 4.1f, tr,
                   &dist);
    }
  // FP64 adjoint
  std::vector<float> coef(7), pos2{nr / 2.f, nc / 2.f, nr / 3.f + 0.5f, nc / 4.f}, d2{0.3f, -0.2f, 1.f, 0.f};
  orc_gauss_coefs(1, 3, 1, coef.data());
  std::vector<double> gd(4), gI(N);
  const double L = orc_lk_loss_grad_f64(f.data(), f.data(), nr, nc, coef.data(), 3, pos2.data(), d2.data(), 2, 7,
                                        gd.data(), gI.data());
  check(std::isfinite(L), "adjoint loss finite");
  // bc2s path
  std::vector<uint8_t> b1(N), b2(N), gi(N);
  std::vector<uint32_t> con(N);
  orc_blur121_u8(v.data(), nr, nc, b1.data());
  orc_blur121_u8(b1.data(), nr, nc, b2.data());
  orc_contrast_u32(v.data(), nr, nc, con.data());
  std::vector<int32_t> taps(7);
  orc_gaussian_int_taps(2.f, 3, taps.data());
  orc_gaussian_int_u8(v.data(), nr, nc, 2.f, 3, gi.data());
  for (int k = 0; k < 40; k++) {
    const int r = (k * 37) % nr, c = (k * 53) % nc;
    uint8_t f16[16];
    orc_bc2s_feature(b1.data(), b2.data(), nr, nc, r, c, f16);
    check(orc_bc2s_distance(b1.data(), b2.data(), nr, nc, f16, r, c, 1) == 0, "bc2s distance at its pixel");
    int32_t d2[2];
    float dist;
    for (int scale = 0; scale < 3; scale++)
      orc_bc2s_match(b1.data(), b2.data(), nr, nc, f16, r + (k % 5) - 2, c - (k % 3) + 1, scale, d2, &dist);
  }
}

// Standalone container routines on a state taken from a tracker.
void container(void* t, int s) {
  int nr, nc;
  orc_tracker_level_dims(t, s, &nr, &nc);
  const int n = orc_tracker_nparticles(t, s);
  std::vector<orc_particle> p(n + 64);
  orc_tracker_get_particles(t, s, p.data(), n);
  std::vector<int32_t> sparse((size_t)nr * nc);
  orc_tracker_get_sparse(t, s, sparse.data());
  std::vector<uint8_t> mask((size_t)nr * nc);
  orc_particle_mask(p.data(), n, nr, nc, mask.data());
  std::vector<orc_particle> q(p);
  std::vector<int32_t> sq(sparse);
  orc_merge(q.data(), n, sq.data(), nr, nc);
  int live = 0;
  for (int i = 0; i < n; i++) live += q[i].age > 0;
  std::vector<uint32_t> matches(n + 1);
  const int kept = orc_particles_compact(q.data(), n, sq.data(), nr, nc, matches.data());
  check(kept == live, "compaction keeps the live particles");
  for (int i = 0; i < kept; i++) {
    const int r = (int)q[i].pos_r, c = (int)q[i].pos_c;
    if (r >= 0 && r < nr && c >= 0 && c < nc && sq[(size_t)r * nc + c] != i) {
      // two particles may share a pixel; the sparse image holds one of them
      const int o = sq[(size_t)r * nc + c];
      check(o >= 0 && o < kept && (int)q[o].pos_r == r && (int)q[o].pos_c == c, "sparse image consistent");
    }
  }
  const int32_t pts[6] = {nr / 2, nc / 2, 4, 4, nr - 5, nc - 5};
  const int m = orc_particles_add(q.data(), kept, kept + 64, sq.data(), nr, nc, pts, 3);
  check(m >= kept && m <= kept + 3, "add");
}

void tracker_runs(const std::vector<uint8_t>& v, int nr, int nc, int nframes) {
  const size_t N = (size_t)nr * nc;
  for (int profile = 0; profile < 3; profile++)
    for (int lanes : {0, 16}) {
      if (profile == 2 && lanes) continue;  // integer strategy: one arithmetic
      orc_tracker_params P = params(profile, 3, 12.f, profile == 0 && lanes ? 64 : lanes);
      if (profile == 2) P.detector_frequency = 3;
      void* a = orc_tracker_create(nr, nc, &P);
      void* b = orc_tracker_create(nr, nc, &P);
      check(orc_tracker_nscales(a) == 3, "nscales");
      for (int f = 0; f < nframes; f++) {
        orc_tracker_run(a, &v[f * N]);
        orc_tracker_update_input(b, &v[f * N]);
        orc_tracker_step(b);
        if (f == nframes / 2) {
          orc_tracker_clear_scale(a, 1);
          orc_tracker_clear_scale(b, 1);
        }
      }
      for (int s = 0; s < 3; s++) {
        int snr, snc;
        orc_tracker_level_dims(a, s, &snr, &snc);
        const int n = orc_tracker_nparticles(a, s);
        check(n == orc_tracker_nparticles(b, s), "update_input + step = run (count)");
        std::vector<orc_particle> pa(n), pb(n);
        orc_tracker_get_particles(a, s, pa.data(), n);
        orc_tracker_get_particles(b, s, pb.data(), n);
        check(n == 0 || std::memcmp(pa.data(), pb.data(), n * sizeof(orc_particle)) == 0,
              "update_input + step = run (particles)");
        std::vector<uint32_t> m(n + 1);
        if (orc_tracker_compact_has_run(a, s)) orc_tracker_get_matches(a, s, m.data(), n + 1);
        const size_t Ns = (size_t)snr * snc;
        std::vector<int32_t> sp(Ns);
        std::vector<uint8_t> sal(Ns), lev(Ns);
        std::vector<float> A(Ns), B(Ns), G(2 * Ns);
        orc_tracker_get_sparse(a, s, sp.data());
        orc_tracker_get_saliency(a, s, sal.data());
        orc_tracker_get_level(a, s, lev.data());
        orc_tracker_get_matcher_images(a, s, A.data(), B.data(), G.data());
        std::vector<int32_t> acc(n + 1);
        std::vector<float> tr(2 * n + 2), dist(n + 1);
        orc_tracker_get_decisions(a, s, acc.data(), tr.data(), dist.data(), n);
        if (profile != 2) orc_tracker_lk_decide(a, s, acc.data(), tr.data(), dist.data(), n);
        if (profile == 2) {
          std::vector<uint8_t> f16(16 * (size_t)(n + 1)), i1(Ns), i2(Ns);
          std::vector<uint32_t> cst(Ns);
          check(orc_tracker_get_features(a, s, f16.data(), n) == n, "features");
          orc_tracker_get_bc2s_images(a, s, i1.data(), i2.data(), cst.data());
          int cr, cc;
          orc_tracker_flow_dims(a, s, &cr, &cc);
          std::vector<float> fl(2 * (size_t)cr * cc);
          std::vector<uint32_t> msc((size_t)cr * cc);
          orc_tracker_get_flow(a, s, fl.data(), msc.data());
        }
        container(a, s);
      }
      // settings of a scale, particles set from outside, clear
      orc_tracker_set_scale(a, 0, 5, 10, 12, nr - 15, nc - 20);
      const int n0 = orc_tracker_nparticles(a, 0);
      std::vector<orc_particle> p0(n0);
      orc_tracker_get_particles(a, 0, p0.data(), n0);
      orc_tracker_set_particles(a, 0, p0.data(), n0);
      orc_tracker_run(a, &v[0]);
      orc_tracker_run(a, &v[N]);
      orc_tracker_clear(a);
      check(orc_tracker_nparticles(a, 0) == 0, "clear empties scale 0");
      orc_tracker_run(a, &v[0]);
      orc_tracker_destroy(a);
      orc_tracker_destroy(b);
    }
}

}  // namespace

int main(int argc, char** argv) {
  const int nframes = argc > 1 ? std::atoi(argv[1]) : 6;
  const int nr = 97, nc = 131;  // odd sizes: floor halving, partial boxes
  std::printf("oracle threads: %d\n", orc_max_threads());
  const std::vector<uint8_t> v = texture_video(0x5A17C0DEull, nr, nc, nframes < 2 ? 2 : nframes);
  images(v, nr, nc);
  std::printf("images done\n");
  {
    // empty containers
    std::vector<int32_t> sp((size_t)nr * nc, -1);
    std::vector<uint8_t> mask((size_t)nr * nc);
    orc_particle none[1];
    uint32_t m[1];
    orc_merge(none, 0, sp.data(), nr, nc);
    orc_particle_mask(none, 0, nr, nc, mask.data());
    check(orc_particles_compact(none, 0, sp.data(), nr, nc, m) == 0, "empty compaction");
    check(orc_particles_add(none, 0, 0, sp.data(), nr, nc, nullptr, 0) == 0, "empty add");
  }
  tracker_runs(v, nr, nc, nframes < 2 ? 2 : nframes);
  std::printf("tracker runs done\n");
  if (failures) {
    std::printf("%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
