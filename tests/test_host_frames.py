"""Host-frame checks of the Python mirror (no device call): the array handed to
the C ABI always has the tracker's shape, contiguous rows and a positive
pitch, so the library's staging copy of nrows rows never reads past it."""
import numpy as np
import pytest

import cuimg_amd as cu


def test_short_and_long_frames_are_rejected():
    with pytest.raises(ValueError):
        cu.host_frame(np.zeros((239, 320), np.uint8), (240, 320))
    with pytest.raises(ValueError):
        cu.host_frame(np.zeros((240, 319), np.uint8), (240, 320))
    with pytest.raises(ValueError):
        cu.host_frame(np.zeros((241, 320), np.uint8), (240, 320))
    with pytest.raises(ValueError):
        cu.host_frame(np.zeros((240, 320, 2), np.uint8), (240, 320))
    with pytest.raises(ValueError):
        cu.host_frame(np.zeros((240, 320), np.float32), (240, 320))


def test_flipped_and_strided_views_become_contiguous_rows():
    a = np.arange(240 * 320, dtype=np.uint32).astype(np.uint8).reshape(240, 320)
    for view in (a[::-1], a[:, ::-1], a[::-1, ::-1]):
        f = cu.host_frame(view, (240, 320))
        assert f.strides[0] > 0 and f.strides[1] == 1
        assert np.array_equal(f, view)
    big = np.zeros((480, 640), np.uint8)
    f = cu.host_frame(big[::2, ::2], (240, 320))
    assert f.strides == (320, 1)
    # a wider host pitch with contiguous rows is passed through as is
    padded = np.zeros((240, 400), np.uint8)
    f = cu.host_frame(padded[:, :320], (240, 320))
    assert f.strides == (400, 1) and np.shares_memory(f, padded)


def test_interleaved_frames():
    c3 = np.zeros((240, 320, 3), np.uint8)
    assert cu.host_frame(c3, (240, 320)).strides == (960, 3, 1)
    f = cu.host_frame(c3[::-1], (240, 320))
    assert f.strides == (960, 3, 1)
    planar = np.zeros((3, 240, 320), np.uint8).transpose(1, 2, 0)  # channels not interleaved
    f = cu.host_frame(planar, (240, 320))
    assert f.strides == (960, 3, 1) and np.array_equal(f, planar)
    one = np.zeros((240, 320, 1), np.uint8)
    assert cu.host_frame(one, (240, 320)).ndim == 2
