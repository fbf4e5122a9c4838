"""Python host mirror of the cuimg tracking hot path on MI355X.

Thin ctypes front end of libcuimg_hip.so (C ABI in include/cuimg_hip.h).  It
mirrors the reference's C++ operator surface for this path — image handles
(host_image2d / device_image2d), the meta-template Gaussian, subsample /
fill_border_clamp, the FAST detector and tracker<pyrlk>(domain, nscales) with
run / scale(s).pset() accessors — so tests and benchmarks read like the
reference's own call sites.  There is no CPU fallback: every operation runs on
the HIP device and raises if the library or the device is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuimg_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "cuimg_hip.h")

U8, F32, F32X4, U32, F32X2 = 1, 2, 3, 4, 5
PYR_BOX2X2, PYR_GAUSS_DECIMATE = 0, 1
_NP = {U8: (np.uint8, ()), F32: (np.float32, ()), F32X4: (np.float32, (4,)),
       U32: (np.uint32, ()), F32X2: (np.float32, (2,))}
_ESIZE = {U8: 1, F32: 4, F32X4: 16, U32: 4, F32X2: 8}

PARTICLE_DTYPE = np.dtype([("speed_r", "<f4"), ("speed_c", "<f4"), ("pos_r", "<f4"),
                           ("pos_c", "<f4"), ("age", "<u2"), ("pad", "<u2")])


class CuimgError(RuntimeError):
    pass


class Img(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_void_p), ("nrows", ctypes.c_int32), ("ncols", ctypes.c_int32),
                ("pitch", ctypes.c_int32), ("border", ctypes.c_int32), ("dtype", ctypes.c_int32)]


class LkParams(ctypes.Structure):
    _fields_ = [("ws", ctypes.c_int32), ("max_iters", ctypes.c_int32), ("eps", ctypes.c_float),
                ("top_n", ctypes.c_int32), ("k", ctypes.c_int32), ("template_u8", ctypes.c_int32)]


class TrackerParams(ctypes.Structure):
    _fields_ = [("nscales", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("lk_ws", ctypes.c_int32), ("lk_max_iters", ctypes.c_int32),
                ("lk_eps", ctypes.c_float), ("lk_top_n", ctypes.c_int32),
                ("fast_n", ctypes.c_int32), ("fast_th", ctypes.c_float),
                ("box_size", ctypes.c_int32), ("detector_frequency", ctypes.c_int32),
                ("filtering_frequency", ctypes.c_int32), ("k", ctypes.c_int32),
                ("with_merge", ctypes.c_int32), ("blur_sigma", ctypes.c_int32),
                ("blur_K", ctypes.c_int32)]


class ScaleParams(ctypes.Structure):
    """cuimg_scale_params: the settings the reference keeps per scale."""
    _fields_ = [("fast_n", ctypes.c_int32), ("fast_th", ctypes.c_float),
                ("detector_frequency", ctypes.c_int32), ("filtering_frequency", ctypes.c_int32),
                ("k", ctypes.c_int32), ("with_merge", ctypes.c_int32), ("box_size", ctypes.c_int32),
                ("pd_r1", ctypes.c_int32), ("pd_c1", ctypes.c_int32), ("pd_r2", ctypes.c_int32),
                ("pd_c2", ctypes.c_int32)]


def profile_params(profile, nscales, fast_th=20.0, **kw):
    """'R': the reference matcher defaults on u8 frames (11x11 window, <=30
    iterations, 25 strongest-gradient pixels); 'B': 7x7 window, <=5 iterations,
    all pixels, on sigma=1 / 7-tap blurred FP32 frames; 'C': the bc2s
    descriptor with the gradient-descent matcher (generic_strategy, the
    reference sample's bc2s_fast_gradient_cpu; the LK settings are unused)."""
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
    p.lk_eps, p.fast_n, p.fast_th, p.box_size = 0.1, 9, fast_th, 3
    p.detector_frequency, p.filtering_frequency, p.k, p.with_merge = 1, 1, 300, 1
    p.blur_sigma, p.blur_K = 1, 3
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_lib = None


def lib():
    """Loads libcuimg_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CuimgError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, c_int, P = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER
        L.cuimg_hip_last_error.restype = ctypes.c_char_p
        L.cuimg_hip_last_error.argtypes = [vp]
        L.cuimg_hip_stream.restype = vp
        L.cuimg_hip_ctx_create.argtypes = [c_int, P(vp)]
        for name in ("cuimg_hip_ctx_destroy", "cuimg_hip_sync", "cuimg_hip_profile_reset"):
            getattr(L, name).argtypes = [vp]
        L.cuimg_hip_profile_enable.argtypes = [vp, c_int]
        L.cuimg_hip_profile_query.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_int64), P(ctypes.c_double)]
        L.cuimg_hip_malloc.argtypes = [vp, ctypes.c_size_t, P(vp)]
        L.cuimg_hip_free.argtypes = [vp, vp]
        L.cuimg_hip_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.cuimg_hip_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.cuimg_hip_memcpy_d2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
        L.cuimg_hip_timer_start.argtypes = [vp]
        L.cuimg_hip_timer_stop.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.cuimg_hip_memset.argtypes = [vp, vp, c_int, ctypes.c_size_t]
        L.cuimg_hip_image_alloc.argtypes = [vp, c_int, c_int, c_int, c_int, P(Img)]
        L.cuimg_hip_image_free.argtypes = [vp, P(Img)]
        L.cuimg_hip_image_upload.argtypes = [vp, P(Img), vp, ctypes.c_size_t]
        L.cuimg_hip_image_download.argtypes = [vp, P(Img), vp, ctypes.c_size_t]
        L.cuimg_hip_fill_border_clamp.argtypes = [vp, P(Img)]
        L.cuimg_hip_gaussian_coefs.argtypes = [c_int, c_int, P(ctypes.c_float)]
        L.cuimg_hip_gaussian_sep.argtypes = [vp, P(Img), P(Img), c_int, c_int]
        L.cuimg_hip_convolve_sep.argtypes = [vp, P(Img), P(Img), P(ctypes.c_float), c_int]
        L.cuimg_hip_convolve_sep_th.argtypes = [vp, P(Img), P(Img), P(ctypes.c_float), c_int, c_int]
        L.cuimg_hip_gaussian_sep_multi.argtypes = [vp, P(Img), P(Img), c_int, c_int, c_int]
        L.cuimg_hip_pyr_down.argtypes = [vp, P(Img), P(Img), c_int]
        L.cuimg_hip_pyramid_build.argtypes = [vp, P(Img), P(Img), c_int]
        L.cuimg_hip_fast_saliency.argtypes = [vp, P(Img), P(Img), P(Img), c_int, c_int]
        L.cuimg_synth_video.argtypes = [ctypes.c_uint64, c_int, c_int, c_int, c_int, vp]
        L.cuimg_hip_tracker_create.argtypes = [vp, c_int, c_int, P(TrackerParams), P(vp)]
        L.cuimg_hip_tracker_destroy.argtypes = [vp]
        L.cuimg_hip_tracker_run.argtypes = [vp, P(Img)]
        L.cuimg_hip_tracker_run_host.argtypes = [vp, vp, ctypes.c_size_t]
        L.cuimg_hip_tracker_run_host_rgb.argtypes = [vp, vp, ctypes.c_size_t, c_int]
        L.cuimg_hip_tracker_clear.argtypes = [vp]
        L.cuimg_hip_tracker_clear_scale.argtypes = [vp, c_int]
        L.cuimg_hip_tracker_update_input.argtypes = [vp, P(Img)]
        L.cuimg_hip_tracker_update_input_host.argtypes = [vp, vp, ctypes.c_size_t]
        L.cuimg_hip_tracker_step.argtypes = [vp]
        L.cuimg_hip_tracker_set_scale_params.argtypes = [vp, c_int, P(ScaleParams)]
        L.cuimg_hip_tracker_get_scale_params.argtypes = [vp, c_int, P(ScaleParams)]
        L.cuimg_hip_tracker_nparticles.argtypes = [vp, c_int, P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_get_particles.argtypes = [vp, c_int, vp, ctypes.c_uint32, P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_set_particles.argtypes = [vp, c_int, vp, ctypes.c_uint32]
        L.cuimg_hip_tracker_get_matches.argtypes = [vp, c_int, vp, ctypes.c_uint32, P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_compact_has_run.argtypes = [vp, c_int, P(c_int)]
        L.cuimg_hip_tracker_sparse.argtypes = [vp, c_int, P(Img)]
        L.cuimg_hip_tracker_level.argtypes = [vp, c_int, P(Img)]
        L.cuimg_hip_tracker_saliency.argtypes = [vp, c_int, P(Img)]
        L.cuimg_hip_tracker_get_decisions.argtypes = [vp, c_int, vp, vp, vp, ctypes.c_uint32, P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_matcher_images.argtypes = [vp, c_int, P(Img), P(Img), P(Img)]
        L.cuimg_hip_tracker_frame_count.argtypes = [vp, P(c_int)]
        L.cuimg_hip_particles_merge.argtypes = [vp, vp, ctypes.c_uint32, P(Img)]
        L.cuimg_hip_particle_mask.argtypes = [vp, vp, ctypes.c_uint32, P(Img)]
        L.cuimg_hip_particles_add.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, ctypes.c_uint32,
                                              P(Img), P(ctypes.c_uint32)]
        L.cuimg_hip_particles_compact.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, P(Img), P(ctypes.c_uint32)]
        L.cuimg_hip_fast_select.argtypes = [vp, P(Img), P(Img), c_int, ctypes.c_float, vp, ctypes.c_uint32,
                                            P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_work.argtypes = [vp, c_int, P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.cuimg_hip_tracker_set_graph.argtypes = [vp, c_int]
        L.cuimg_hip_tracker_set_tuning.argtypes = [vp, c_int, c_int]
        L.cuimg_hip_lk_track.argtypes = [vp, P(Img), P(Img), c_int, c_int, vp, ctypes.c_uint32, P(LkParams),
                                         P(Img), vp, vp, vp]
        L.cuimg_hip_lk_track_noalloc.argtypes = L.cuimg_hip_lk_track.argtypes
        L.cuimg_hip_tracker_get_features.argtypes = [vp, c_int, vp, ctypes.c_uint32, P(ctypes.c_uint32)]
        L.cuimg_hip_tracker_bc2s_images.argtypes = [vp, c_int, P(Img), P(Img), P(Img)]
        L.cuimg_hip_tracker_flow.argtypes = [vp, c_int, P(Img), P(Img)]
        L.cuimg_hip_bc2s_images.argtypes = [vp, P(Img), P(Img), P(Img), P(Img)]
        L.cuimg_hip_blur121.argtypes = [vp, P(Img), P(Img)]
        L.cuimg_hip_gaussian_int.argtypes = [vp, P(Img), P(Img), ctypes.c_float, c_int]
        L.cuimg_hip_bc2s_features.argtypes = [vp, P(Img), P(Img), vp, ctypes.c_uint32, vp]
        L.cuimg_hip_bc2s_match.argtypes = [vp, P(Img), P(Img), vp, vp, ctypes.c_uint32, c_int, vp, vp]
        L.cuimg_hip_batch_create.argtypes = [vp, c_int, P(c_int), P(c_int), P(TrackerParams), P(vp)]
        L.cuimg_hip_batch_destroy.argtypes = [vp]
        L.cuimg_hip_batch_step.argtypes = [vp, P(Img)]
        L.cuimg_hip_batch_step_host.argtypes = [vp, P(vp), P(ctypes.c_size_t), c_int]
        L.cuimg_hip_batch_set_graph.argtypes = [vp, c_int]
        L.cuimg_hip_batch_tracker.argtypes = [vp, c_int, P(vp)]
        L.cuimg_hip_batch_set_tuning.argtypes = [vp, c_int, c_int]
        L.cuimg_hip_batch_get_tuning.argtypes = [vp, P(c_int), P(c_int)]
        L.cuimg_hip_batch_clear.argtypes = [vp]
        L.cuimg_hip_lk_loss_adjoint.argtypes = [vp, P(Img), P(Img), vp, vp, ctypes.c_uint32, c_int, c_int, c_int,
                                                vp, vp, vp, P(Img), vp, ctypes.c_size_t]
        _lib = L
    return _lib


def synth_video(seed, nrows, ncols, nframes, motion):
    """Seeded synthetic texture video (host code of the library)."""
    out = np.zeros((nframes, nrows, ncols), np.uint8)
    r = lib().cuimg_synth_video(seed, nrows, ncols, nframes, motion, out.ctypes.data)
    if r:
        raise CuimgError(f"cuimg_synth_video failed ({r})")
    return out


def gaussian_coefs(sigma, K):
    out = np.zeros(2 * K + 1, np.float32)
    r = lib().cuimg_hip_gaussian_coefs(sigma, K, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if r:
        raise CuimgError("gaussian_coefs failed")
    return out


class Context:
    """One HIP device + stream (cuimg_hip_ctx)."""

    def __init__(self, device=0):
        L = lib()
        self.L = L
        h = ctypes.c_void_p()
        r = L.cuimg_hip_ctx_create(device, ctypes.byref(h))
        if r:
            raise CuimgError(f"cuimg_hip_ctx_create({device}) failed with {r} (no HIP device?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.cuimg_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, r, what):
        if r:
            msg = self.L.cuimg_hip_last_error(self.h)
            raise CuimgError(f"{what}: error {r}: {msg.decode() if msg else ''}")

    def sync(self):
        self.check(self.L.cuimg_hip_sync(self.h), "sync")

    # profiling
    def profile(self, on=True):
        self.check(self.L.cuimg_hip_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_reset(self):
        self.check(self.L.cuimg_hip_profile_reset(self.h), "profile_reset")

    def profile_query(self, name):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        self.check(self.L.cuimg_hip_profile_query(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)),
                   "profile_query")
        return n.value, ms.value

    def timer_start(self):
        """Records the start event of a span on the context stream."""
        self.check(self.L.cuimg_hip_timer_start(self.h), "timer_start")

    def timer_stop(self):
        """Milliseconds of the context stream's work since timer_start."""
        ms = ctypes.c_double()
        self.check(self.L.cuimg_hip_timer_stop(self.h, ctypes.byref(ms)), "timer_stop")
        return ms.value

    # images
    def image(self, nrows, ncols, dtype, border=0):
        return DeviceImage(self, nrows, ncols, dtype, border)

    def upload(self, arr, border=0):
        arr = np.ascontiguousarray(arr)
        if arr.dtype == np.uint8:
            dt = U8
        elif arr.dtype == np.float32 and arr.ndim == 2:
            dt = F32
        elif arr.dtype == np.float32 and arr.ndim == 3 and arr.shape[2] == 4:
            dt = F32X4
        else:
            raise ValueError(f"unsupported array {arr.dtype} {arr.shape}")
        im = self.image(arr.shape[0], arr.shape[1], dt, border)
        im.upload(arr)
        return im

    def gaussian(self, src, dst, sigma, K):
        self.check(self.L.cuimg_hip_gaussian_sep(self.h, ctypes.byref(src.d), ctypes.byref(dst.d), sigma, K),
                   "gaussian_sep")

    def convolve(self, src, dst, coefs, K, th=0):
        c = np.ascontiguousarray(coefs, np.float32)
        self.check(self.L.cuimg_hip_convolve_sep_th(self.h, ctypes.byref(src.d), ctypes.byref(dst.d),
                                                    c.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), K, th),
                   "convolve_sep")

    def gaussian_multi(self, srcs, dsts, sigma, K):
        a = (Img * len(srcs))(*[s.d for s in srcs])
        b = (Img * len(dsts))(*[d.d for d in dsts])
        self.check(self.L.cuimg_hip_gaussian_sep_multi(self.h, a, b, len(srcs), sigma, K), "gaussian_sep_multi")

    def pyr_down(self, src, dst, mode=0):
        """One pyramid level: mode PYR_BOX2X2 (subsample) or PYR_GAUSS_DECIMATE."""
        self.check(self.L.cuimg_hip_pyr_down(self.h, ctypes.byref(src.d), ctypes.byref(dst.d), mode), "pyr_down")

    def pyramid(self, frame, levels):
        a = (Img * len(levels))(*[l.d for l in levels])
        self.check(self.L.cuimg_hip_pyramid_build(self.h, ctypes.byref(frame.d), a, len(levels)),
                   "pyramid_build")

    def fill_border(self, img):
        self.check(self.L.cuimg_hip_fill_border_clamp(self.h, ctypes.byref(img.d)), "fill_border_clamp")

    # ---------------------------------------------------------------- bc2s
    def bc2s_images(self, src, s1, s2, contrast=None):
        """s1 = blur121(src), s2 = blur121(s1) (borders clamp-filled), contrast
        (U32, U8 saturated, or None)."""
        c = ctypes.byref(contrast.d) if contrast is not None else None
        self.check(self.L.cuimg_hip_bc2s_images(self.h, ctypes.byref(src.d), ctypes.byref(s1.d), ctypes.byref(s2.d),
                                                c), "bc2s_images")

    def blur121(self, src, dst):
        self.check(self.L.cuimg_hip_blur121(self.h, ctypes.byref(src.d), ctypes.byref(dst.d)), "blur121")

    def gaussian_int(self, src, dst, sigma, hw):
        self.check(self.L.cuimg_hip_gaussian_int(self.h, ctypes.byref(src.d), ctypes.byref(dst.d),
                                                 ctypes.c_float(sigma), int(hw)), "gaussian_int")

    def _dev(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        self.check(self.L.cuimg_hip_malloc(self.h, max(arr.nbytes, 16), ctypes.byref(p)), "malloc")
        if arr.nbytes:
            self.check(self.L.cuimg_hip_memcpy_h2d(self.h, p, arr.ctypes.data, arr.nbytes), "memcpy_h2d")
        return p

    def _host(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        if out.nbytes:
            self.check(self.L.cuimg_hip_memcpy_d2h(self.h, out.ctypes.data, p, out.nbytes), "memcpy_d2h")
        return out

    def bc2s_features(self, s1, s2, pts):
        """Descriptors (n x 16 u8) at (row, col) points of the feature images."""
        pts = np.ascontiguousarray(pts, np.int32).reshape(-1, 2)
        n = len(pts)
        dp, do = self._dev(pts), self._dev(np.zeros((n, 16), np.uint8))
        try:
            self.check(self.L.cuimg_hip_bc2s_features(self.h, ctypes.byref(s1.d), ctypes.byref(s2.d), dp, n, do),
                       "bc2s_features")
            return self._host(do, (n, 16), np.uint8)
        finally:
            self.L.cuimg_hip_free(self.h, dp)
            self.L.cuimg_hip_free(self.h, do)

    def bc2s_match(self, s1, s2, feats, preds, scale=0):
        """Two-step descent (scale 0) or one descent at scale 1 / 2: (n x 2
        int32 displacements, n distances)."""
        feats = np.ascontiguousarray(feats, np.uint8).reshape(-1, 16)
        preds = np.ascontiguousarray(preds, np.int32).reshape(-1, 2)
        n = len(preds)
        assert len(feats) == n
        df, dp = self._dev(feats), self._dev(preds)
        dd, ds = self._dev(np.zeros((n, 2), np.int32)), self._dev(np.zeros(n, np.float32))
        try:
            self.check(self.L.cuimg_hip_bc2s_match(self.h, ctypes.byref(s1.d), ctypes.byref(s2.d), df, dp, n, scale,
                                                   dd, ds), "bc2s_match")
            return self._host(dd, (n, 2), np.int32), self._host(ds, n, np.float32)
        finally:
            for q in (df, dp, dd, ds):
                self.L.cuimg_hip_free(self.h, q)

    def fast(self, src, mask, dst, n=9, th=20):
        m = ctypes.byref(mask.d) if mask is not None else None
        self.check(self.L.cuimg_hip_fast_saliency(self.h, ctypes.byref(src.d), m, ctypes.byref(dst.d), n, th),
                   "fast_saliency")


class DeviceImage:
    """device_image2d<V>(domain, border): owned device image."""

    def __init__(self, ctx, nrows, ncols, dtype, border=0):
        self.ctx = ctx
        self.d = Img()
        ctx.check(ctx.L.cuimg_hip_image_alloc(ctx.h, nrows, ncols, border, dtype, ctypes.byref(self.d)),
                  "image_alloc")

    @property
    def shape(self):
        return (self.d.nrows, self.d.ncols)

    def free(self):
        if self.d.begin and self.ctx.h:
            self.ctx.L.cuimg_hip_image_free(self.ctx.h, ctypes.byref(self.d))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr):
        np_t, tail = _NP[self.d.dtype]
        arr = np.ascontiguousarray(arr, np_t)
        assert arr.shape == (self.d.nrows, self.d.ncols) + tail, (arr.shape, self.shape)
        pitch = self.d.ncols * _ESIZE[self.d.dtype]
        self.ctx.check(self.ctx.L.cuimg_hip_image_upload(self.ctx.h, ctypes.byref(self.d), arr.ctypes.data, pitch),
                       "image_upload")

    def download(self):
        np_t, tail = _NP[self.d.dtype]
        out = np.zeros((self.d.nrows, self.d.ncols) + tail, np_t)
        pitch = self.d.ncols * _ESIZE[self.d.dtype]
        self.ctx.check(self.ctx.L.cuimg_hip_image_download(self.ctx.h, ctypes.byref(self.d), out.ctypes.data, pitch),
                       "image_download")
        return out

    def download_with_border(self):
        """Domain plus border as one array (for border checks)."""
        b = self.d.border
        np_t, tail = _NP[self.d.dtype]
        es = _ESIZE[self.d.dtype]
        nr, nc = self.d.nrows + 2 * b, self.d.ncols + 2 * b
        out = np.zeros((nr, nc) + tail, np_t)
        src = Img(self.d.begin - b * self.d.pitch - b * es, nr, nc, self.d.pitch, 0, self.d.dtype)
        self.ctx.check(self.ctx.L.cuimg_hip_image_download(self.ctx.h, ctypes.byref(src), out.ctypes.data, nc * es),
                       "image_download")
        return out


class View:
    """Non-owning device image view (kernel_image2d) returned by the tracker."""

    def __init__(self, ctx, d):
        self.ctx = ctx
        self.d = d

    @property
    def shape(self):
        return (self.d.nrows, self.d.ncols)

    def download(self):
        np_t, tail = _NP[self.d.dtype]
        out = np.zeros((self.d.nrows, self.d.ncols) + tail, np_t)
        pitch = self.d.ncols * _ESIZE[self.d.dtype]
        self.ctx.check(self.ctx.L.cuimg_hip_image_download(self.ctx.h, ctypes.byref(self.d), out.ctypes.data, pitch),
                       "image_download")
        return out


class Tracker:
    """tracker<pyrlk>(domain, nscales) on the GPU (cuimg_hip_tracker).

    run(frame) is tracker::run; particles(s) / matches(s) / sparse(s) are
    scale(s).pset().dense_particles() / matches() / sparse_particles().  A
    tracker made by Batch.tracker(i) is a view of one stream of a batch."""

    def __init__(self, ctx, nrows, ncols, params, _view=None):
        self.ctx = ctx
        self.L = ctx.L
        self.params = params
        self.nscales = params.nscales
        self.shape = (nrows, ncols)
        if _view is not None:
            self.h, self._owner = _view, False
            return
        h = ctypes.c_void_p()
        ctx.check(self.L.cuimg_hip_tracker_create(ctx.h, nrows, ncols, ctypes.byref(params), ctypes.byref(h)),
                  "tracker_create")
        self.h, self._owner = h, True

    def close(self):
        if getattr(self, "h", None):
            # closing the context destroyed the tracker already: it is only
            # forgotten then
            if self._owner and self.ctx.h:
                self.L.cuimg_hip_tracker_destroy(self.h)
            self.h = None

    def set_graph(self, on):
        """Replay frame steps from a captured hipGraph (default) or launch them."""
        self.ctx.check(self.L.cuimg_hip_tracker_set_graph(self.h, 1 if on else 0), "set_graph")

    def set_tuning(self, lk_ppw, blur_th=0):
        """Launch shape: particles per matcher wave, blur tile rows (0 = automatic)."""
        self.ctx.check(self.L.cuimg_hip_tracker_set_tuning(self.h, lk_ppw, blur_th), "set_tuning")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, frame):
        if isinstance(frame, (DeviceImage, View)):
            self.ctx.check(self.L.cuimg_hip_tracker_run(self.h, ctypes.byref(frame.d)), "tracker_run")
        else:
            frame = host_frame(frame, self.shape)
            if frame.ndim == 3:  # interleaved channels: channel 0 is the gray frame (get_x)
                self.ctx.check(self.L.cuimg_hip_tracker_run_host_rgb(self.h, frame.ctypes.data, frame.strides[0],
                                                                     frame.shape[2]), "tracker_run_host_rgb")
                return
            self.ctx.check(self.L.cuimg_hip_tracker_run_host(self.h, frame.ctypes.data, frame.strides[0]),
                           "tracker_run_host")

    def clear(self):
        self.ctx.check(self.L.cuimg_hip_tracker_clear(self.h), "tracker_clear")

    def update_input(self, frame):
        """tracker::update_input: the pyramid of the next frame (device image or
        host array), without the frame step."""
        if isinstance(frame, (DeviceImage, View)):
            self.ctx.check(self.L.cuimg_hip_tracker_update_input(self.h, ctypes.byref(frame.d)), "update_input")
            return
        frame = host_frame(frame, self.shape)
        if frame.ndim != 2:
            raise ValueError("update_input takes a gray frame")
        self.ctx.check(self.L.cuimg_hip_tracker_update_input_host(self.h, frame.ctypes.data, frame.strides[0]),
                       "update_input")

    def step(self):
        """tracker::run(): the frame step on the input last given to update_input."""
        self.ctx.check(self.L.cuimg_hip_tracker_step(self.h), "tracker_step")

    def clear_scale(self, s):
        """The strategy's clear() of scale s: its frame counter restarts."""
        self.ctx.check(self.L.cuimg_hip_tracker_clear_scale(self.h, s), "clear_scale")

    def scale_params(self, s):
        p = ScaleParams()
        self.ctx.check(self.L.cuimg_hip_tracker_get_scale_params(self.h, s, ctypes.byref(p)), "get_scale_params")
        return p

    def set_scale_params(self, s, p):
        self.ctx.check(self.L.cuimg_hip_tracker_set_scale_params(self.h, s, ctypes.byref(p)), "set_scale_params")

    def nparticles(self, s):
        n = ctypes.c_uint32()
        self.ctx.check(self.L.cuimg_hip_tracker_nparticles(self.h, s, ctypes.byref(n)), "nparticles")
        return n.value

    def particles(self, s):
        n = ctypes.c_uint32()
        self.ctx.check(self.L.cuimg_hip_tracker_get_particles(self.h, s, None, 0, ctypes.byref(n)), "get_particles")
        out = np.zeros(n.value, PARTICLE_DTYPE)
        if n.value:
            self.ctx.check(self.L.cuimg_hip_tracker_get_particles(self.h, s, out.ctypes.data, n.value,
                                                                  ctypes.byref(n)), "get_particles")
        return out

    def set_particles(self, s, parts):
        parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
        self.ctx.check(self.L.cuimg_hip_tracker_set_particles(self.h, s, parts.ctypes.data, len(parts)),
                       "set_particles")

    def matches(self, s):
        n = ctypes.c_uint32()
        self.ctx.check(self.L.cuimg_hip_tracker_get_matches(self.h, s, None, 0, ctypes.byref(n)), "get_matches")
        out = np.zeros(n.value, np.uint32)
        if n.value:
            self.ctx.check(self.L.cuimg_hip_tracker_get_matches(self.h, s, out.ctypes.data, n.value,
                                                                ctypes.byref(n)), "get_matches")
        return out

    def compact_has_run(self, s):
        v = ctypes.c_int()
        self.ctx.check(self.L.cuimg_hip_tracker_compact_has_run(self.h, s, ctypes.byref(v)), "compact_has_run")
        return bool(v.value)

    def _view(self, fn, s):
        d = Img()
        self.ctx.check(fn(self.h, s, ctypes.byref(d)), "tracker view")
        return View(self.ctx, d)

    def sparse(self, s):
        return self._view(self.L.cuimg_hip_tracker_sparse, s).download().view(np.int32)

    def level(self, s):
        return self._view(self.L.cuimg_hip_tracker_level, s).download()

    def saliency(self, s):
        return self._view(self.L.cuimg_hip_tracker_saliency, s).download()

    def matcher_images(self, s):
        a, b, g = Img(), Img(), Img()
        self.ctx.check(self.L.cuimg_hip_tracker_matcher_images(self.h, s, ctypes.byref(a), ctypes.byref(b),
                                                               ctypes.byref(g)), "matcher_images")
        return View(self.ctx, a).download(), View(self.ctx, b).download()

    def decisions(self, s):
        n = ctypes.c_uint32()
        self.ctx.check(self.L.cuimg_hip_tracker_get_decisions(self.h, s, None, None, None, 0, ctypes.byref(n)),
                       "get_decisions")
        acc = np.zeros(n.value, np.int32)
        tr = np.zeros((n.value, 2), np.float32)
        dist = np.zeros(n.value, np.float32)
        if n.value:
            self.ctx.check(self.L.cuimg_hip_tracker_get_decisions(self.h, s, acc.ctypes.data, tr.ctypes.data,
                                                                  dist.ctypes.data, n.value, ctypes.byref(n)),
                           "get_decisions")
        return acc, tr, dist

    def features(self, s):
        """bc2s descriptors of the dense vector (n x 16 u8), profile 'C'."""
        n = ctypes.c_uint32()
        self.ctx.check(self.L.cuimg_hip_tracker_get_features(self.h, s, None, 0, ctypes.byref(n)), "get_features")
        out = np.zeros((n.value, 16), np.uint8)
        if n.value:
            self.ctx.check(self.L.cuimg_hip_tracker_get_features(self.h, s, out.ctypes.data, n.value,
                                                                 ctypes.byref(n)), "get_features")
        return out

    def bc2s_images(self, s):
        """(s1, s2, contrast saturated at 255) of the scale's last frame."""
        a, b, c = Img(), Img(), Img()
        self.ctx.check(self.L.cuimg_hip_tracker_bc2s_images(self.h, s, ctypes.byref(a), ctypes.byref(b),
                                                            ctypes.byref(c)), "bc2s_images")
        return View(self.ctx, a).download(), View(self.ctx, b).download(), View(self.ctx, c).download()

    def flow(self, s):
        """(flow (cells x 2 float32, -500 = none), multiscale count) of the last frame."""
        a, b = Img(), Img()
        self.ctx.check(self.L.cuimg_hip_tracker_flow(self.h, s, ctypes.byref(a), ctypes.byref(b)), "flow")
        return View(self.ctx, a).download(), View(self.ctx, b).download()

    def work(self, s):
        """(particles run through the matcher, accepted) at scale s since creation."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.L.cuimg_hip_tracker_work(self.h, s, ctypes.byref(a), ctypes.byref(b)), "tracker_work")
        return a.value, b.value

    @property
    def frame_count(self):
        v = ctypes.c_int()
        self.ctx.check(self.L.cuimg_hip_tracker_frame_count(self.h, ctypes.byref(v)), "frame_count")
        return v.value


def host_frame(frame, shape):
    """A host frame as the C ABI reads it: u8 rows of (nrows, ncols) pixels
    (2-D) or of interleaved channels (3-D, 1/3/4 channels), each row contiguous
    and rows at a positive pitch.  Anything else (a short or long frame, a
    flipped or strided view) is copied or rejected here, never over-read."""
    frame = np.asarray(frame)
    if frame.dtype != np.uint8:
        raise ValueError(f"host frame must be uint8, not {frame.dtype}")
    if frame.ndim not in (2, 3) or tuple(frame.shape[:2]) != tuple(shape):
        raise ValueError(f"host frame of shape {frame.shape} for a {shape[0]}x{shape[1]} tracker")
    if frame.ndim == 3 and frame.shape[2] not in (1, 3, 4):
        raise ValueError(f"host frame with {frame.shape[2]} channels (1, 3 or 4)")
    if frame.ndim == 3 and frame.shape[2] == 1:
        frame = frame[:, :, 0]
    inner = frame.shape[2] if frame.ndim == 3 else 1
    row_contiguous = frame.strides[1] == inner and (frame.ndim == 2 or frame.strides[2] == 1)
    if frame.strides[0] < frame.shape[1] * inner or not row_contiguous:
        frame = np.ascontiguousarray(frame)
    return frame


class Batch:
    """Independent trackers stepped together (cuimg_hip_batch): shapes is a list
    of (nrows, ncols), params a TrackerParams or one per stream."""

    def __init__(self, ctx, shapes, params):
        self.ctx = ctx
        self.L = ctx.L
        n = len(shapes)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        self.params = plist
        rows = (ctypes.c_int * n)(*[r for r, _ in shapes])
        cols = (ctypes.c_int * n)(*[c for _, c in shapes])
        parr = (TrackerParams * n)(*plist)
        h = ctypes.c_void_p()
        ctx.check(self.L.cuimg_hip_batch_create(ctx.h, n, rows, cols, parr, ctypes.byref(h)), "batch_create")
        self.h = h
        self.n = n
        self.shapes = list(shapes)
        self._views = {}

    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:  # see Tracker.close
                self.L.cuimg_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, frames):
        """One frame step of every stream: device images, or host arrays
        (2-D gray, or 3-D interleaved channels with channel 0 the gray frame;
        every stream the same channel count)."""
        if all(isinstance(f, (DeviceImage, View)) for f in frames):
            arr = (Img * self.n)(*[f.d for f in frames])
            self.ctx.check(self.L.cuimg_hip_batch_step(self.h, arr), "batch_step")
            return
        if len(frames) != self.n:
            raise ValueError(f"{len(frames)} frames for {self.n} streams")
        hs = [host_frame(f, shape) for f, shape in zip(frames, self.shapes)]
        ch = hs[0].shape[2] if hs[0].ndim == 3 else 1
        if any((f.shape[2] if f.ndim == 3 else 1) != ch for f in hs):
            raise ValueError("every stream's host frame needs the same channel count")
        ptrs = (ctypes.c_void_p * self.n)(*[f.ctypes.data for f in hs])
        pitches = (ctypes.c_size_t * self.n)(*[f.strides[0] for f in hs])
        self.ctx.check(self.L.cuimg_hip_batch_step_host(self.h, ptrs, pitches, ch), "batch_step_host")

    def set_graph(self, on):
        self.ctx.check(self.L.cuimg_hip_batch_set_graph(self.h, 1 if on else 0), "batch_set_graph")

    def set_tuning(self, lk_ppw, blur_th=0):
        self.ctx.check(self.L.cuimg_hip_batch_set_tuning(self.h, lk_ppw, blur_th), "batch_set_tuning")

    def tuning(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.L.cuimg_hip_batch_get_tuning(self.h, ctypes.byref(a), ctypes.byref(b)), "batch_get_tuning")
        return a.value, b.value

    def clear(self):
        self.ctx.check(self.L.cuimg_hip_batch_clear(self.h), "batch_clear")

    def autotune(self, run, candidates=((8, 0), (4, 0), (16, 0), (32, 0), (8, 16), (8, 32), (8, 64))):
        """Launch-shape autotuner: for each (lk_ppw, blur_th) candidate the
        batch is cleared and `run(batch)` is called; it steps the batch over a
        representative sequence and returns the seconds it measured.  The
        fastest candidate is kept; returns (best, {candidate: seconds})."""
        results = {}
        for c in candidates:
            self.set_tuning(*c)
            self.clear()
            results[c] = run(self)
        best = min(results, key=results.get)
        self.set_tuning(*best)
        self.clear()
        return best, results

    def tracker(self, i):
        if i not in self._views:
            h = ctypes.c_void_p()
            self.ctx.check(self.L.cuimg_hip_batch_tracker(self.h, i, ctypes.byref(h)), "batch_tracker")
            r, c = self.shapes[i]
            self._views[i] = Tracker(self.ctx, r, c, self.params[i], _view=h)
        return self._views[i]


def lk_track(ctx, prev_levels, cur_levels, scale, parts, ws=7, max_iters=5, eps=0.1, top_n=49, k=300,
             template_u8=0, sparse=None):
    """pyrlk_match_particles_kernel for the particles of one scale: returns
    (particles after the move, accept, translation, dist)."""
    L = ctx.L
    n = len(parts)
    parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
    prm = LkParams(ws, max_iters, eps, top_n, k, template_u8)
    a = (Img * len(prev_levels))(*[x.d for x in prev_levels])
    b = (Img * len(cur_levels))(*[x.d for x in cur_levels])
    buf = DeviceBuffer.from_array(ctx, parts) if n else None
    acc, tr, dist = (DeviceBuffer(ctx, max(4 * n, 4)), DeviceBuffer(ctx, max(8 * n, 8)),
                     DeviceBuffer(ctx, max(4 * n, 4)))
    try:
        ctx.check(L.cuimg_hip_lk_track(ctx.h, a, b, len(prev_levels), scale, buf.p if buf else None, n,
                                       ctypes.byref(prm), ctypes.byref(sparse.d) if sparse is not None else None,
                                       acc.p, tr.p, dist.p), "lk_track")
        out = buf.download(PARTICLE_DTYPE, n) if n else parts.copy()
        return (out, acc.download(np.int32, n), tr.download(np.float32, 2 * n).reshape(n, 2),
                dist.download(np.float32, n))
    finally:
        for x in (buf, acc, tr, dist):
            if x is not None:
                x.free()


def particles_merge(ctx, parts, sparse):
    """merge_trajectories over a container state on the device: returns the
    particle vector after the merge.  `sparse` is the (nrows, ncols) index image
    (-1 = empty)."""
    L = ctx.L
    parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
    sp = np.ascontiguousarray(sparse).astype(np.int32).view(np.uint32)
    img = ctx.image(sp.shape[0], sp.shape[1], U32)
    img.upload(sp)
    dev = ctypes.c_void_p()
    nbytes = max(parts.nbytes, 1)
    ctx.check(L.cuimg_hip_malloc(ctx.h, nbytes, ctypes.byref(dev)), "malloc")
    try:
        if len(parts):
            ctx.check(L.cuimg_hip_memcpy_h2d(ctx.h, dev, parts.ctypes.data, parts.nbytes), "h2d")
        ctx.check(L.cuimg_hip_particles_merge(ctx.h, dev, len(parts), ctypes.byref(img.d)), "particles_merge")
        out = np.zeros_like(parts)
        if len(parts):
            ctx.check(L.cuimg_hip_memcpy_d2h(ctx.h, out.ctypes.data, dev, parts.nbytes), "d2h")
    finally:
        L.cuimg_hip_free(ctx.h, dev)
    return out


def particle_mask(ctx, parts, nrows, ncols):
    """Detector mask of a dense particle vector on the device: (nrows, ncols)
    uint8, 255 except 0 on the 3x3 block of every particle."""
    L = ctx.L
    parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
    img = ctx.image(nrows, ncols, U8)
    dev = DeviceBuffer(ctx, parts.nbytes)
    if len(parts):
        ctx.check(L.cuimg_hip_memcpy_h2d(ctx.h, dev.p, parts.ctypes.data, parts.nbytes), "h2d")
    ctx.check(L.cuimg_hip_particle_mask(ctx.h, dev.p, len(parts), ctypes.byref(img.d)), "particle_mask")
    out = img.download()
    dev.free()
    return out


def fast_select(ctx, sal, sparse, box, th, cap=None):
    """New-point selection on the device: int32 (n, 2) array of (row, col) in
    raster order.  `sparse` is the (nrows, ncols) index image (-1 = empty)."""
    L = ctx.L
    sal = np.ascontiguousarray(sal, np.uint8)
    sp = np.ascontiguousarray(sparse).astype(np.int32).view(np.uint32)
    nr, nc = sal.shape
    s_img = ctx.image(nr, nc, U8)
    s_img.upload(sal)
    p_img = ctx.image(nr, nc, U32)
    p_img.upload(sp)
    if cap is None:
        cap = max(1, (nr // max(1, box)) * (nc // max(1, box)))
    dev = DeviceBuffer(ctx, 8 * cap)
    n = ctypes.c_uint32()
    ctx.check(L.cuimg_hip_fast_select(ctx.h, ctypes.byref(s_img.d), ctypes.byref(p_img.d), box, float(th),
                                      dev.p, cap, ctypes.byref(n)), "fast_select")
    k = min(n.value, cap)
    out = np.zeros((k, 2), np.int32)
    if k:
        ctx.check(L.cuimg_hip_memcpy_d2h(ctx.h, out.ctypes.data, dev.p, out.nbytes), "d2h")
    dev.free()
    return out, n.value


def _sparse_image(ctx, sparse):
    sp = np.ascontiguousarray(sparse).astype(np.int32).view(np.uint32)
    img = ctx.image(sp.shape[0], sp.shape[1], U32)
    img.upload(sp)
    return img


def particles_add(ctx, parts, sparse, pts, cap=None):
    """particle_container::add of each (row, col) of pts on the device while
    below cap: returns (particles, sparse) with sparse as int32 (-1 = empty)."""
    L = ctx.L
    parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
    pts = np.ascontiguousarray(np.asarray(pts, np.int32).reshape(-1, 2))
    cap = len(parts) + len(pts) if cap is None else cap
    img = _sparse_image(ctx, sparse)
    dev = DeviceBuffer(ctx, max(cap, len(parts)) * PARTICLE_DTYPE.itemsize)
    if len(parts):
        ctx.check(L.cuimg_hip_memcpy_h2d(ctx.h, dev.p, parts.ctypes.data, parts.nbytes), "h2d")
    dpts = DeviceBuffer.from_array(ctx, pts)
    n = ctypes.c_uint32()
    ctx.check(L.cuimg_hip_particles_add(ctx.h, dev.p, len(parts), cap, dpts.p, len(pts), ctypes.byref(img.d),
                                        ctypes.byref(n)), "particles_add")
    out = dev.download(PARTICLE_DTYPE, (n.value,))
    sp = img.download().view(np.int32)
    dev.free()
    dpts.free()
    return out, sp


def particles_compact(ctx, parts, sparse):
    """particle_container::compact on the device: returns (particles, matches,
    sparse) with matches as uint32 (0xFFFFFFFF = removed) and sparse as int32."""
    L = ctx.L
    parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
    img = _sparse_image(ctx, sparse)
    src = DeviceBuffer.from_array(ctx, parts)
    dst = DeviceBuffer(ctx, parts.nbytes)
    mt = DeviceBuffer(ctx, 4 * len(parts))
    n = ctypes.c_uint32()
    ctx.check(L.cuimg_hip_particles_compact(ctx.h, src.p, len(parts), dst.p, mt.p, ctypes.byref(img.d),
                                            ctypes.byref(n)), "particles_compact")
    out = dst.download(PARTICLE_DTYPE, (n.value,))
    matches = mt.download(np.uint32, (len(parts),))
    sp = img.download().view(np.int32)
    for b in (src, dst, mt):
        b.free()
    return out, matches, sp


class DeviceBuffer:
    """Plain device allocation (cuimg_hip_malloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.p = ctypes.c_void_p()
        self.nbytes = max(int(nbytes), 1)
        ctx.check(ctx.L.cuimg_hip_malloc(ctx.h, self.nbytes, ctypes.byref(self.p)), "malloc")

    @classmethod
    def from_array(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(ctx, arr.nbytes)
        if arr.nbytes:
            ctx.check(ctx.L.cuimg_hip_memcpy_h2d(ctx.h, b.p, arr.ctypes.data, arr.nbytes), "h2d")
        return b

    def copy_from(self, src, nbytes=None):
        """Device-to-device copy of nbytes (default: all) from another buffer."""
        n = self.nbytes if nbytes is None else nbytes
        self.ctx.check(self.ctx.L.cuimg_hip_memcpy_d2d(self.ctx.h, self.p, src.p, n), "memcpy_d2d")

    def download(self, dtype, shape):
        out = np.zeros(shape, dtype)
        if out.nbytes:
            self.ctx.check(self.ctx.L.cuimg_hip_memcpy_d2h(self.ctx.h, out.ctypes.data, self.p, out.nbytes), "d2h")
        return out

    def free(self):
        if self.p and self.ctx.h:
            self.ctx.L.cuimg_hip_free(self.ctx.h, self.p)
            self.p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def lk_loss_adjoint(ctx, I0, I1, pos, d, sigma=1, K=3, ws=7, want_dI=True):
    """Photometric LK loss and gradients on the device.  I0, I1: DeviceImage
    (F32); pos, d: (n, 2) float32 host arrays.  Returns (L, loss_p, dL_dd,
    dL_dI1 or None) as host arrays."""
    L = ctx.L
    n = len(pos)
    bp = DeviceBuffer.from_array(ctx, np.asarray(pos, np.float32))
    bd = DeviceBuffer.from_array(ctx, np.asarray(d, np.float32))
    tot = DeviceBuffer(ctx, 8)
    lp = DeviceBuffer(ctx, 4 * n)
    gd = DeviceBuffer(ctx, 8 * n)
    nr, nc = I1.shape
    scratch = DeviceBuffer(ctx, 8 * nr * nc)
    gI = ctx.image(nr, nc, F32) if want_dI else None
    ctx.check(L.cuimg_hip_lk_loss_adjoint(ctx.h, ctypes.byref(I0.d), ctypes.byref(I1.d), bp.p, bd.p, n, sigma, K,
                                          ws, tot.p, lp.p, gd.p, ctypes.byref(gI.d) if gI else None, scratch.p,
                                          2 * nr * nc), "lk_loss_adjoint")
    ctx.sync()
    return (float(tot.download(np.float64, (1,))[0]), lp.download(np.float32, (n,)),
            gd.download(np.float32, (n, 2)), gI.download() if gI else None)
