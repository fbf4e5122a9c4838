"""ctypes front end of the CPU oracle (liboracle_cuimg.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a C++ restatement of the reference
tracking hot path (each function cites the reference file:line it follows).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
and only as the checker or as the timed CPU baseline; the product never does.

Two arithmetic readings where the reference's rounding is not reproducible
(it adds window terms in loop order with whatever contraction its compiler
chose): the default rounds every product before adding it, in the reference's
loop order; the "GPU arithmetic" mode (tracker sum_lanes 16 / 64, gaussian
fused=True) performs the GPU's operations: fused multiply-adds for the blur
taps and the last three bilinear terms, window sums in lane-group order.

Parity status of the restated functions (what pins each one):
  * Gaussian coefficients: the reference's own generator, built from its source
    under oracle/_ref, plus a committed fixture of its output;
  * separable Gaussian: impulse / constant / ramp closed forms, brute-force 2D
    sums, an FP64 twin; the fused mode by a Python restatement with exactly
    rounded FMAs;
  * box pyramid: integer closed form;
  * FAST saliency: known-answer corners and an independent Python restatement;
  * LK matcher: bilinear exactness, zero motion, recovered translations,
    degenerate-G failure, rejection branches, and bit-exact agreement with an
    independent numpy restatement in both arithmetic readings;
  * detector mask and box selection: independent Python restatements
    (Chebyshev-distance mask, numpy first-occurrence argmax per box);
  * particle container add / compact / merge: independent Python restatements
    of the serial semantics and the reference's own assert after compaction;
  * coarse-to-fine recursion: a known large translation recovered on 4 levels
    and not on one;
  * frame step and container moves: the static-video known answer; on an
    exact integer-translation video with detector_frequency 3 and
    filtering_frequency 2, the closed forms of the detection / compaction /
    merge cadence, ages, positions (birth pixel + age x shift) and the sparse
    image's last writer; crafted merges whose outcome follows by hand from the
    reference's merge loop and c8 order; committed end-to-end fixtures of
    the 320x240 video (regression pin);
  * the photometric loss gradient: central finite differences.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CUIMG_ORACLE_LIB selects another build of the same library (the
# UndefinedBehaviorSanitizer build of `make -C oracle ubsan-lib`).
LIB_PATH = os.environ.get("CUIMG_ORACLE_LIB") or os.path.join(_HERE, "liboracle_cuimg.so")


class Particle(ctypes.Structure):
    _fields_ = [("speed_r", ctypes.c_float), ("speed_c", ctypes.c_float),
                ("pos_r", ctypes.c_float), ("pos_c", ctypes.c_float),
                ("age", ctypes.c_uint16), ("pad", ctypes.c_uint16)]


PARTICLE_DTYPE = np.dtype([("speed_r", "<f4"), ("speed_c", "<f4"), ("pos_r", "<f4"),
                           ("pos_c", "<f4"), ("age", "<u2"), ("pad", "<u2")])
assert PARTICLE_DTYPE.itemsize == 20


class TrackerParams(ctypes.Structure):
    _fields_ = [("nscales", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("lk_ws", ctypes.c_int32), ("lk_max_iters", ctypes.c_int32),
                ("lk_eps", ctypes.c_float), ("lk_top_n", ctypes.c_int32),
                ("fast_n", ctypes.c_int32), ("fast_th", ctypes.c_float),
                ("box_size", ctypes.c_int32), ("detector_frequency", ctypes.c_int32),
                ("filtering_frequency", ctypes.c_int32), ("k", ctypes.c_int32),
                ("with_merge", ctypes.c_int32), ("blur_sigma", ctypes.c_int32),
                ("blur_K", ctypes.c_int32), ("sum_lanes", ctypes.c_int32)]


def gpu_sum_lanes(ws):
    """Lanes per particle of the GPU matcher for a window side: the oracle's
    window sums in that group order reproduce the GPU's rounding."""
    return 16 if ws <= 7 else 64


def profile_params(profile, nscales, fast_th=20.0, **kw):
    """Profile 'R' = the reference's matcher defaults on u8 frames
    (tracking2/pyrlk_tracker.hpp:173,212,283); profile 'B' = 7x7 window,
    <=5 iterations, on sigma=1 / 7-tap blurred FP32 frames; profile 'C' = the
    bc2s descriptor with the gradient-descent matcher (generic_strategy,
    the reference sample's bc2s_fast_gradient_cpu; LK settings unused)."""
    p = TrackerParams()
    p.nscales = nscales
    if profile == "R":
        p.profile, p.lk_ws, p.lk_max_iters, p.lk_top_n = 0, 11, 30, 25
    elif profile == "B":
        p.profile, p.lk_ws, p.lk_max_iters, p.lk_top_n = 1, 7, 5, 49
    elif profile == "C":
        p.profile, p.lk_ws, p.lk_max_iters, p.lk_top_n = 2, 7, 5, 49
    else:
        raise ValueError(profile)
    p.lk_eps = 0.1
    p.fast_n = 9
    p.fast_th = fast_th
    p.box_size = 3
    p.detector_frequency = 1
    p.filtering_frequency = 1
    p.k = 300
    p.with_merge = 1
    p.blur_sigma = 1
    p.blur_K = 3
    p.sum_lanes = 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        self.L = L
        c_int, c_float, c_double, vp = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
        P = ctypes.POINTER
        L.orc_gauss_coef_runtime.restype = c_float
        L.orc_gauss_coef_runtime.argtypes = [c_float, c_int]
        L.orc_gauss_coef_meta.restype = c_float
        L.orc_gauss_coef_meta.argtypes = [c_int, c_int]
        L.orc_gauss_coefs.argtypes = [c_int, c_int, c_int, P(c_float)]
        L.orc_gaussian_sep_f32.argtypes = [P(c_float), c_int, c_int, c_int, P(c_float), c_int, c_int, P(c_float)]
        L.orc_gaussian_sep_u8.argtypes = [P(ctypes.c_uint8), c_int, c_int, P(c_float), c_int, c_int, P(c_float)]
        L.orc_gaussian_sep_f64.argtypes = [P(c_double), c_int, c_int, c_int, P(c_double), c_int, P(c_double)]
        L.orc_pyr_down_u8.argtypes = [P(ctypes.c_uint8), c_int, c_int, P(ctypes.c_uint8)]
        L.orc_pyr_down_f32.argtypes = [P(c_float), c_int, c_int, P(c_float)]
        L.orc_fast_score.restype = ctypes.c_uint8
        L.orc_fast_score.argtypes = [P(ctypes.c_uint8), c_int, c_int, c_int, c_int, c_int, c_int]
        L.orc_fast_saliency.argtypes = [P(ctypes.c_uint8), c_int, c_int, P(ctypes.c_uint8), c_int, c_int,
                                        P(ctypes.c_uint8)]
        L.orc_sobel_gradient.argtypes = [P(c_float), c_int, c_int, c_int, P(c_float)]
        L.orc_lk_match.argtypes = [P(c_float), P(c_float), P(c_float), c_int, c_int, c_int, c_int,
                                   c_float, c_int, c_int, c_int, c_float, c_float, c_float, c_float,
                                   P(c_float), P(c_float)]
        L.orc_tracker_create.restype = vp
        L.orc_tracker_create.argtypes = [c_int, c_int, P(TrackerParams)]
        for name in ("orc_tracker_destroy", "orc_tracker_clear"):
            getattr(L, name).argtypes = [vp]
        L.orc_tracker_run.argtypes = [vp, P(ctypes.c_uint8)]
        L.orc_tracker_update_input.argtypes = [vp, P(ctypes.c_uint8)]
        L.orc_tracker_step.argtypes = [vp]
        L.orc_tracker_clear_scale.argtypes = [vp, c_int]
        L.orc_tracker_set_scale.argtypes = [vp, c_int, c_int, c_int, c_int, c_int, c_int]
        L.orc_tracker_nscales.argtypes = [vp]
        L.orc_tracker_level_dims.argtypes = [vp, c_int, P(c_int), P(c_int)]
        L.orc_tracker_nparticles.argtypes = [vp, c_int]
        L.orc_tracker_get_particles.argtypes = [vp, c_int, vp, c_int]
        L.orc_tracker_set_particles.argtypes = [vp, c_int, vp, c_int]
        L.orc_tracker_get_matches.argtypes = [vp, c_int, P(ctypes.c_uint32), c_int]
        L.orc_tracker_compact_has_run.argtypes = [vp, c_int]
        L.orc_tracker_get_sparse.argtypes = [vp, c_int, P(ctypes.c_int32)]
        L.orc_tracker_get_saliency.argtypes = [vp, c_int, P(ctypes.c_uint8)]
        L.orc_tracker_get_level.argtypes = [vp, c_int, P(ctypes.c_uint8)]
        L.orc_tracker_get_matcher_images.argtypes = [vp, c_int, vp, vp, vp]
        L.orc_tracker_get_decisions.argtypes = [vp, c_int, P(ctypes.c_int32), P(c_float), P(c_float), c_int]
        L.orc_tracker_lk_decide.argtypes = [vp, c_int, P(ctypes.c_int32), P(c_float), P(c_float), c_int]
        L.orc_merge.argtypes = [vp, c_int, P(ctypes.c_int32), c_int, c_int]
        L.orc_particle_mask.argtypes = [vp, c_int, c_int, c_int, P(ctypes.c_uint8)]
        L.orc_fast_select.restype = c_int
        L.orc_set_threads.argtypes = [c_int]
        L.orc_max_threads.restype = c_int
        L.orc_particles_add.restype = c_int
        L.orc_particles_add.argtypes = [vp, c_int, c_int, P(ctypes.c_int32), c_int, c_int, P(ctypes.c_int32), c_int]
        L.orc_particles_compact.restype = c_int
        L.orc_particles_compact.argtypes = [vp, c_int, P(ctypes.c_int32), c_int, c_int, P(ctypes.c_uint32)]
        L.orc_fast_select.argtypes = [P(ctypes.c_uint8), P(ctypes.c_int32), c_int, c_int, c_int, c_float,
                                      P(ctypes.c_int32), c_int]
        u8p, i32p, u32p = P(ctypes.c_uint8), P(ctypes.c_int32), P(ctypes.c_uint32)
        L.orc_blur121_u8.argtypes = [u8p, c_int, c_int, u8p]
        L.orc_gaussian_int_taps.argtypes = [c_float, c_int, i32p]
        L.orc_gaussian_int_u8.argtypes = [u8p, c_int, c_int, c_float, c_int, u8p]
        L.orc_contrast_u32.argtypes = [u8p, c_int, c_int, u32p]
        L.orc_bc2s_feature.argtypes = [u8p, u8p, c_int, c_int, c_int, c_int, u8p]
        L.orc_bc2s_distance.restype = c_int
        L.orc_bc2s_distance.argtypes = [u8p, u8p, c_int, c_int, u8p, c_int, c_int, c_int]
        L.orc_bc2s_match.argtypes = [u8p, u8p, c_int, c_int, u8p, c_int, c_int, c_int, i32p, P(c_float)]
        L.orc_tracker_get_features.restype = c_int
        L.orc_tracker_get_features.argtypes = [vp, c_int, u8p, c_int]
        L.orc_tracker_get_bc2s_images.argtypes = [vp, c_int, vp, vp, vp]
        L.orc_tracker_flow_dims.argtypes = [vp, c_int, P(c_int), P(c_int)]
        L.orc_tracker_get_flow.argtypes = [vp, c_int, vp, vp]
        L.orc_lk_loss_grad_f64.restype = c_double
        L.orc_lk_loss_grad_f64.argtypes = [P(c_float), P(c_float), c_int, c_int, P(c_float), c_int,
                                           P(c_float), P(c_float), c_int, c_int, vp, vp]
        L.orc_lk_loss_grad_f64_d.restype = c_double
        L.orc_lk_loss_grad_f64_d.argtypes = [P(c_double), P(c_double), c_int, c_int, P(c_double), c_int,
                                             P(c_double), P(c_double), c_int, c_int, vp, vp]

    # ---------------------------------------------------------------- gaussian
    def coef_runtime(self, sigma, x):
        return self.L.orc_gauss_coef_runtime(float(sigma), int(x))

    def coef_meta(self, sigma, x):
        return self.L.orc_gauss_coef_meta(int(sigma), int(x))

    def coefs(self, sigma, K, meta=True):
        out = np.zeros(2 * K + 1, np.float32)
        self.L.orc_gauss_coefs(int(sigma), int(K), 1 if meta else 0, _ptr(out, ctypes.c_float))
        return out

    def gaussian(self, img, coef, K, fused=False):
        """fused=True: taps added with fused multiply-adds (the GPU's arithmetic)."""
        coef = np.ascontiguousarray(coef, np.float32)
        f = 1 if fused else 0
        if img.dtype == np.uint8:
            img = np.ascontiguousarray(img)
            out = np.zeros(img.shape, np.float32)
            self.L.orc_gaussian_sep_u8(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1],
                                       _ptr(coef, ctypes.c_float), K, f, _ptr(out, ctypes.c_float))
            return out
        img = np.ascontiguousarray(img, np.float32)
        ch = 1 if img.ndim == 2 else img.shape[2]
        out = np.zeros(img.shape, np.float32)
        self.L.orc_gaussian_sep_f32(_ptr(img, ctypes.c_float), img.shape[0], img.shape[1], ch,
                                    _ptr(coef, ctypes.c_float), K, f, _ptr(out, ctypes.c_float))
        return out

    def gaussian_f64(self, img, coef, K):
        img = np.ascontiguousarray(img, np.float64)
        coef = np.ascontiguousarray(coef, np.float64)
        ch = 1 if img.ndim == 2 else img.shape[2]
        out = np.zeros(img.shape, np.float64)
        self.L.orc_gaussian_sep_f64(_ptr(img, ctypes.c_double), img.shape[0], img.shape[1], ch,
                                    _ptr(coef, ctypes.c_double), K, _ptr(out, ctypes.c_double))
        return out

    # ---------------------------------------------------------------- pyramid
    def pyr_down(self, img):
        img = np.ascontiguousarray(img)
        nr, nc = img.shape
        out = np.zeros((nr // 2, nc // 2), img.dtype)
        if img.dtype == np.uint8:
            self.L.orc_pyr_down_u8(_ptr(img, ctypes.c_uint8), nr, nc, _ptr(out, ctypes.c_uint8))
        else:
            self.L.orc_pyr_down_f32(_ptr(img, ctypes.c_float), nr, nc, _ptr(out, ctypes.c_float))
        return out

    def pyr_down_gauss(self, img):
        """build_mipmap's blur-then-subsample level (cuimg/gpu/mipmap.h:94-109):
        the sigma-1, K-6 Gaussian of the image sampled at 2p, FP32."""
        nr, nc = img.shape
        g = self.gaussian(img, self.coefs(1, 6), 6)
        return np.ascontiguousarray(g[0:2 * (nr // 2):2, 0:2 * (nc // 2):2])

    # ---------------------------------------------------------------- FAST
    def fast_score(self, img, r, c, n=9, th=20):
        img = np.ascontiguousarray(img, np.uint8)
        return self.L.orc_fast_score(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1], r, c, n, th)

    def fast_saliency(self, img, mask=None, n=9, th=20):
        img = np.ascontiguousarray(img, np.uint8)
        if mask is None:
            mask = np.full(img.shape, 255, np.uint8)
        mask = np.ascontiguousarray(mask, np.uint8)
        out = np.zeros(img.shape, np.uint8)
        self.L.orc_fast_saliency(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1],
                                 _ptr(mask, ctypes.c_uint8), n, th, _ptr(out, ctypes.c_uint8))
        return out

    # ---------------------------------------------------------------- LK
    def sobel(self, prev, integer_source):
        prev = np.ascontiguousarray(prev, np.float32)
        out = np.zeros(prev.shape + (2,), np.float32)
        self.L.orc_sobel_gradient(_ptr(prev, ctypes.c_float), prev.shape[0], prev.shape[1],
                                  1 if integer_source else 0, _ptr(out, ctypes.c_float))
        return out

    def lk_match(self, A, B, G, p, pred, ws=7, max_iters=5, eps=0.1, top_n=49, template_u8=0, sum_lanes=0):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        G = np.ascontiguousarray(G, np.float32)
        tr = np.zeros(2, np.float32)
        dist = np.zeros(1, np.float32)
        self.L.orc_lk_match(_ptr(A, ctypes.c_float), _ptr(B, ctypes.c_float), _ptr(G, ctypes.c_float),
                            A.shape[0], A.shape[1], ws, max_iters, eps, top_n, template_u8, sum_lanes,
                            float(p[0]), float(p[1]), float(pred[0]), float(pred[1]),
                            _ptr(tr, ctypes.c_float), _ptr(dist, ctypes.c_float))
        return tr, float(dist[0])

    def merge(self, parts, sparse):
        parts = np.array(parts, PARTICLE_DTYPE, copy=True)
        sparse = np.ascontiguousarray(sparse, np.int32)
        self.L.orc_merge(parts.ctypes.data, len(parts), _ptr(sparse, ctypes.c_int32), sparse.shape[0],
                         sparse.shape[1])
        return parts

    def particle_mask(self, parts, nr, nc):
        """Detector mask of a dense particle vector (uint8 nr x nc)."""
        parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
        mask = np.zeros((nr, nc), np.uint8)
        self.L.orc_particle_mask(parts.ctypes.data if len(parts) else None, len(parts), nr, nc,
                                 _ptr(mask, ctypes.c_uint8))
        return mask

    def fast_select(self, sal, sparse, box, th):
        """New-point selection: int32 (n, 2) array of (row, col) in raster order."""
        sal = np.ascontiguousarray(sal, np.uint8)
        sparse = np.ascontiguousarray(sparse, np.int32)
        nr, nc = sal.shape
        cap = max(1, (nr // max(1, box) + 1) * (nc // max(1, box) + 1))
        pts = np.zeros((cap, 2), np.int32)
        n = self.L.orc_fast_select(_ptr(sal, ctypes.c_uint8), _ptr(sparse, ctypes.c_int32), nr, nc, box,
                                   float(th), _ptr(pts, ctypes.c_int32), cap)
        return pts[:min(n, cap)].copy()

    def set_threads(self, n):
        """OpenMP threads of the oracle's parallel loops."""
        self.L.orc_set_threads(int(n))

    def max_threads(self):
        return int(self.L.orc_max_threads())

    def particles_add(self, parts, sparse, pts, cap=None):
        """particle_container::add for each (row, col) of pts while below cap:
        returns (particles, sparse)."""
        parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
        pts = np.ascontiguousarray(np.asarray(pts, np.int32).reshape(-1, 2))
        cap = len(parts) + len(pts) if cap is None else cap
        buf = np.zeros(max(cap, len(parts), 1), PARTICLE_DTYPE)
        buf[:len(parts)] = parts
        sp = np.array(sparse, np.int32, copy=True)
        n = self.L.orc_particles_add(buf.ctypes.data, len(parts), cap, _ptr(sp, ctypes.c_int32), sp.shape[0],
                                     sp.shape[1], _ptr(pts, ctypes.c_int32), len(pts))
        return buf[:n].copy(), sp

    def particles_compact(self, parts, sparse):
        """particle_container::compact(cpu): returns (particles, matches, sparse)."""
        buf = np.array(parts, PARTICLE_DTYPE, copy=True)
        sp = np.array(sparse, np.int32, copy=True)
        m = np.zeros(max(len(buf), 1), np.uint32)
        n = self.L.orc_particles_compact(buf.ctypes.data if len(buf) else None, len(buf),
                                         _ptr(sp, ctypes.c_int32), sp.shape[0], sp.shape[1],
                                         _ptr(m, ctypes.c_uint32))
        return buf[:n].copy(), m[:len(buf)].copy(), sp

    # ---------------------------------------------------------------- tracker
    def tracker(self, nr, nc, params):
        return OracleTracker(self, nr, nc, params)

    # ---------------------------------------------------------------- bc2s
    def blur121(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros_like(img)
        self.L.orc_blur121_u8(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1], _ptr(out, ctypes.c_uint8))
        return out

    def gaussian_int_taps(self, sigma, hw):
        out = np.zeros(2 * hw + 1, np.int32)
        self.L.orc_gaussian_int_taps(ctypes.c_float(sigma), int(hw), _ptr(out, ctypes.c_int32))
        return out

    def gaussian_int(self, img, sigma, hw):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros_like(img)
        self.L.orc_gaussian_int_u8(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1], ctypes.c_float(sigma),
                                   int(hw), _ptr(out, ctypes.c_uint8))
        return out

    def contrast(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros(img.shape, np.uint32)
        self.L.orc_contrast_u32(_ptr(img, ctypes.c_uint8), img.shape[0], img.shape[1], _ptr(out, ctypes.c_uint32))
        return out

    def bc2s_images(self, img):
        """(s1, s2): the bc2s feature images of a u8 frame."""
        s1 = self.blur121(img)
        return s1, self.blur121(s1)

    def bc2s_feature(self, s1, s2, r, c):
        s1 = np.ascontiguousarray(s1, np.uint8)
        s2 = np.ascontiguousarray(s2, np.uint8)
        out = np.zeros(16, np.uint8)
        self.L.orc_bc2s_feature(_ptr(s1, ctypes.c_uint8), _ptr(s2, ctypes.c_uint8), s1.shape[0], s1.shape[1],
                                int(r), int(c), _ptr(out, ctypes.c_uint8))
        return out

    def bc2s_distance(self, s1, s2, f, r, c, scale=1):
        s1 = np.ascontiguousarray(s1, np.uint8)
        s2 = np.ascontiguousarray(s2, np.uint8)
        f = np.ascontiguousarray(f, np.uint8)
        return self.L.orc_bc2s_distance(_ptr(s1, ctypes.c_uint8), _ptr(s2, ctypes.c_uint8), s1.shape[0],
                                        s1.shape[1], _ptr(f, ctypes.c_uint8), int(r), int(c), int(scale))

    def bc2s_match(self, s1, s2, f, pred, scale=0):
        """scale 0: two_step_gradient_descent_match; 1 / 2: one descent.
        Returns ((dr, dc), distance)."""
        s1 = np.ascontiguousarray(s1, np.uint8)
        s2 = np.ascontiguousarray(s2, np.uint8)
        f = np.ascontiguousarray(f, np.uint8)
        d = np.zeros(2, np.int32)
        dist = ctypes.c_float()
        self.L.orc_bc2s_match(_ptr(s1, ctypes.c_uint8), _ptr(s2, ctypes.c_uint8), s1.shape[0], s1.shape[1],
                              _ptr(f, ctypes.c_uint8), int(pred[0]), int(pred[1]), int(scale),
                              _ptr(d, ctypes.c_int32), ctypes.byref(dist))
        return (int(d[0]), int(d[1])), float(dist.value)

    # ---------------------------------------------------------------- adjoint
    def loss_grad(self, I0, I1, coef, K, pos, d, ws=7, want_dI=True):
        I0 = np.ascontiguousarray(I0, np.float32)
        I1 = np.ascontiguousarray(I1, np.float32)
        coef = np.ascontiguousarray(coef, np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        n = pos.shape[0]
        gd = np.zeros((n, 2), np.float64)
        gI = np.zeros(I1.shape, np.float64) if want_dI else None
        L = self.L.orc_lk_loss_grad_f64(_ptr(I0, ctypes.c_float), _ptr(I1, ctypes.c_float),
                                        I0.shape[0], I0.shape[1], _ptr(coef, ctypes.c_float), K,
                                        _ptr(pos, ctypes.c_float), _ptr(d, ctypes.c_float), n, ws,
                                        gd.ctypes.data, gI.ctypes.data if want_dI else None)
        return L, gd, gI

    def loss_f64(self, I0, I1, coef, K, pos, d, ws=7):
        I0 = np.ascontiguousarray(I0, np.float64)
        I1 = np.ascontiguousarray(I1, np.float64)
        coef = np.ascontiguousarray(coef, np.float64)
        pos = np.ascontiguousarray(pos, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        return self.L.orc_lk_loss_grad_f64_d(_ptr(I0, ctypes.c_double), _ptr(I1, ctypes.c_double),
                                             I0.shape[0], I0.shape[1], _ptr(coef, ctypes.c_double),
                                             K, _ptr(pos, ctypes.c_double), _ptr(d, ctypes.c_double),
                                             pos.shape[0], ws, None, None)


class OracleTracker:
    def __init__(self, orc, nr, nc, params):
        self.o = orc
        self.L = orc.L
        self.params = params
        self.h = self.L.orc_tracker_create(nr, nc, ctypes.byref(params))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_tracker_destroy(self.h)
            self.h = None

    def run(self, frame):
        frame = np.ascontiguousarray(frame, np.uint8)
        self.L.orc_tracker_run(self.h, _ptr(frame, ctypes.c_uint8))

    def clear(self):
        self.L.orc_tracker_clear(self.h)

    def update_input(self, frame):
        frame = np.ascontiguousarray(frame, np.uint8)
        self.L.orc_tracker_update_input(self.h, _ptr(frame, ctypes.c_uint8))

    def step(self):
        self.L.orc_tracker_step(self.h)

    def clear_scale(self, s):
        self.L.orc_tracker_clear_scale(self.h, s)

    def set_scale(self, s, box_size, domain):
        """fast_detector::set_box_size and pyrlk_cpu::set_particle_domain of
        scale s; domain = (r1, c1, r2, c2) inclusive."""
        self.L.orc_tracker_set_scale(self.h, s, box_size, *domain)

    @property
    def nscales(self):
        return self.L.orc_tracker_nscales(self.h)

    def dims(self, s):
        nr, nc = ctypes.c_int(), ctypes.c_int()
        self.L.orc_tracker_level_dims(self.h, s, ctypes.byref(nr), ctypes.byref(nc))
        return nr.value, nc.value

    def particles(self, s):
        n = self.L.orc_tracker_nparticles(self.h, s)
        out = np.zeros(n, PARTICLE_DTYPE)
        if n:
            self.L.orc_tracker_get_particles(self.h, s, out.ctypes.data, n)
        return out

    def set_particles(self, s, parts):
        parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
        self.L.orc_tracker_set_particles(self.h, s, parts.ctypes.data, len(parts))

    def matches(self, s):
        cap = self.L.orc_tracker_get_matches(self.h, s, None, 0)
        n = -cap if cap < 0 else 0
        out = np.zeros(n, np.uint32)
        if n:
            self.L.orc_tracker_get_matches(self.h, s, _ptr(out, ctypes.c_uint32), n)
        return out

    def compact_has_run(self, s):
        return bool(self.L.orc_tracker_compact_has_run(self.h, s))

    def sparse(self, s):
        nr, nc = self.dims(s)
        out = np.zeros((nr, nc), np.int32)
        self.L.orc_tracker_get_sparse(self.h, s, _ptr(out, ctypes.c_int32))
        return out

    def saliency(self, s):
        nr, nc = self.dims(s)
        out = np.zeros((nr, nc), np.uint8)
        self.L.orc_tracker_get_saliency(self.h, s, _ptr(out, ctypes.c_uint8))
        return out

    def level(self, s):
        nr, nc = self.dims(s)
        out = np.zeros((nr, nc), np.uint8)
        self.L.orc_tracker_get_level(self.h, s, _ptr(out, ctypes.c_uint8))
        return out

    def matcher_images(self, s):
        nr, nc = self.dims(s)
        A = np.zeros((nr, nc), np.float32)
        B = np.zeros((nr, nc), np.float32)
        G = np.zeros((nr, nc, 2), np.float32)
        self.L.orc_tracker_get_matcher_images(self.h, s, A.ctypes.data, B.ctypes.data, G.ctypes.data)
        return A, B, G

    def features(self, s):
        """bc2s descriptors of the dense vector (n x 16 bytes)."""
        n = self.L.orc_tracker_nparticles(self.h, s)
        out = np.zeros((max(n, 1), 16), np.uint8)
        k = self.L.orc_tracker_get_features(self.h, s, _ptr(out, ctypes.c_uint8), n)
        assert k == n, (k, n)
        return out[:n]

    def bc2s_images(self, s):
        """(s1, s2, contrast) of the scale's last frame."""
        nr, nc = self.dims(s)
        s1 = np.zeros((nr, nc), np.uint8)
        s2 = np.zeros((nr, nc), np.uint8)
        con = np.zeros((nr, nc), np.uint32)
        self.L.orc_tracker_get_bc2s_images(self.h, s, s1.ctypes.data, s2.ctypes.data, con.ctypes.data)
        return s1, s2, con

    def flow(self, s):
        """(flow (cr, cc, 2) float32, multiscale count (cr, cc) uint32)."""
        cr, cc = ctypes.c_int(), ctypes.c_int()
        self.L.orc_tracker_flow_dims(self.h, s, ctypes.byref(cr), ctypes.byref(cc))
        f = np.zeros((cr.value, cc.value, 2), np.float32)
        m = np.zeros((cr.value, cc.value), np.uint32)
        self.L.orc_tracker_get_flow(self.h, s, f.ctypes.data, m.ctypes.data)
        return f, m

    def _decisions(self, fn, s):
        n = self.L.orc_tracker_nparticles(self.h, s) + 1
        while True:
            acc = np.zeros(n, np.int32)
            tr = np.zeros((n, 2), np.float32)
            dist = np.zeros(n, np.float32)
            k = fn(self.h, s, _ptr(acc, ctypes.c_int32), _ptr(tr, ctypes.c_float),
                   _ptr(dist, ctypes.c_float), n)
            if k >= 0:
                return acc[:k], tr[:k], dist[:k]
            n = -k

    def last_decisions(self, s):
        return self._decisions(self.L.orc_tracker_get_decisions, s)

    def lk_decide(self, s):
        return self._decisions(self.L.orc_tracker_lk_decide, s)
