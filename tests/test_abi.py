"""CPU checks of the C-ABI library: it loads, exports every function the public
header declares, and its host-only entry points (coefficients, synthetic
video) behave; no device call is made."""
import ctypes
import os
import re

import numpy as np

import cuimg_amd as cu

HEADER = cu.HEADER


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cuimg_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(cu.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_header_is_plain_c():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for bad in ("torch", "std::", "hip_runtime"):
        assert bad not in text
    assert not re.search(r"\btemplate\s*<", text)


def test_gaussian_coefs_match_oracle(orc):
    for sigma in (1, 2, 4):
        K = 3 * sigma
        assert np.array_equal(cu.gaussian_coefs(sigma, K), orc.coefs(sigma, K))


def test_synth_video_deterministic_and_moving():
    a = cu.synth_video(0xC0FFEE01, 96, 128, 4, 0)
    b = cu.synth_video(0xC0FFEE01, 96, 128, 4, 0)
    assert np.array_equal(a, b)
    c = cu.synth_video(0xC0FFEE02, 96, 128, 4, 0)
    assert not np.array_equal(a, c)
    assert a[0].std() > 20
    # consecutive frames differ (motion + noise) but stay correlated
    d = np.abs(a[1].astype(int) - a[0].astype(int)).mean()
    assert 1 < d < 60


def test_no_device_is_reported_not_crashing():
    # on a host without a GPU the context creation fails cleanly
    if os.path.exists("/dev/kfd"):
        return
    try:
        cu.Context(0)
    except cu.CuimgError:
        return
    raise AssertionError("Context() succeeded without a device")


def test_entry_points_reject_null_handles():
    """Every stateful entry point checks its handle before any device call:
    NULL contexts, trackers and batches give an error code, not a crash."""
    L = cu.lib()
    img = cu.Img()
    n = ctypes.c_uint32()
    calls = [
        lambda: L.cuimg_hip_particle_mask(None, None, 0, ctypes.byref(img)),
        lambda: L.cuimg_hip_fast_select(None, ctypes.byref(img), ctypes.byref(img), 3, 20.0, None, 0, ctypes.byref(n)),
        lambda: L.cuimg_hip_particles_add(None, None, 0, 0, None, 0, ctypes.byref(img), ctypes.byref(n)),
        lambda: L.cuimg_hip_particles_compact(None, None, 0, None, None, ctypes.byref(img), ctypes.byref(n)),
        lambda: L.cuimg_hip_particles_merge(None, None, 0, ctypes.byref(img)),
        lambda: L.cuimg_hip_tracker_run_host(None, None, 0),
        lambda: L.cuimg_hip_tracker_run_host_rgb(None, None, 0, 3),
        lambda: L.cuimg_hip_batch_step_host(None, None, None, 1),
        lambda: L.cuimg_hip_pyr_down(None, None, None, 0),
        lambda: L.cuimg_hip_lk_loss_adjoint(None, ctypes.byref(img), ctypes.byref(img), None, None, 0, 1, 3, 7, None,
                                            None, None, None, None, 0),
    ]
    for call in calls:
        assert call() != 0
