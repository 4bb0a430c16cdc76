"""Pinning the oracle's frame pyramid (GSSlam::initAndCopyImgs, GSSlam.cu:1913-1969) on the CPU.

cv::pyrDown is not importable here (no OpenCV), so its documented definition is restated with
scipy: a 5x5 convolution with [1 4 6 4 1]^T [1 4 6 4 1] / 256 in scipy's "mirror" mode (=
OpenCV's BORDER_REFLECT_101), keeping even rows and columns. The 8-bit path must equal
floor(x + 1/2) of that exactly (integer sums); the float path within FP32 rounding. Closed forms
pin the size rule, constants and the 8-bit / 16-bit conversions.
"""
import numpy as np
import pytest
from scipy.ndimage import convolve

import oracle as O

K5 = np.array([1, 4, 6, 4, 1], np.float64)
K55 = np.outer(K5, K5) / 256.0


def _scipy_pyrdown(img):
    img = np.asarray(img, np.float64)
    if img.ndim == 3:
        return np.stack([_scipy_pyrdown(img[..., c]) for c in range(img.shape[2])], axis=-1)
    return convolve(img, K55, mode="mirror")[::2, ::2]


@pytest.mark.parametrize("h,w", [(5, 7), (48, 64), (479, 641), (2, 2), (1, 9), (3, 3)])
def test_pyrdown_u8_equals_documented_kernel_rounded_half_up(h, w):
    rng = np.random.default_rng(h * 1000 + w)
    img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    got = O.pyrdown_u8(img)
    assert got.shape == ((h + 1) // 2, (w + 1) // 2, 3)
    want = np.floor(_scipy_pyrdown(img) + 0.5).astype(np.uint8)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("h,w", [(5, 7), (480, 640), (479, 641)])
def test_pyrdown_f32_matches_documented_kernel(h, w):
    rng = np.random.default_rng(h + w)
    img = rng.uniform(0.3, 6.0, (h, w)).astype(np.float32)
    got = O.pyrdown_f32(img)
    want = _scipy_pyrdown(img)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=0)


def test_pyrdown_constants_and_ramps():
    np.testing.assert_array_equal(O.pyrdown_u8(np.full((9, 11, 3), 200, np.uint8)), 200)
    np.testing.assert_array_equal(O.pyrdown_f32(np.full((9, 11), 1.5, np.float32)), 1.5)
    # a linear ramp is reproduced inside (the kernel is symmetric with unit sum) ...
    x = np.tile(np.arange(16, dtype=np.float32), (8, 1))
    d = O.pyrdown_f32(x)
    np.testing.assert_array_equal(d[:, 1:-1], x[::2, ::2][:, 1:-1])
    # ... and reflect-101 at the border: column 0 sees (2, 1, 0, 1, 2) -> (1*2+4*1+6*0+4*1+1*2)/16
    np.testing.assert_allclose(d[:, 0], 12.0 / 16.0)


def test_conversions_and_pyramid_levels():
    rng = np.random.default_rng(3)
    col = rng.integers(0, 256, (61, 83, 3), dtype=np.uint8)
    col[0, 0] = (0, 128, 255)
    d16 = rng.integers(0, 65535, (61, 83), dtype=np.uint16)
    levels = O.frame_pyramid(col, d16, 3)
    assert [lv[0].shape[:2] for lv in levels] == [(61, 83), (31, 42), (16, 21)]
    rgb0, dep0, gx0, gy0 = levels[0]
    assert rgb0[0, 0, 0] == 0.0 and rgb0[0, 0, 2] == 1.0
    np.testing.assert_array_equal(rgb0, col.astype(np.float32) * np.float32(1 / 255))
    np.testing.assert_array_equal(dep0, d16.astype(np.float32) * np.float32(0.001))
    # level 1 = pyrDown of the 8-bit colour, then converted (not pyrDown of the float colour)
    np.testing.assert_array_equal(levels[1][0],
                                  O.pyrdown_u8(col).astype(np.float32) * np.float32(1 / 255))
    np.testing.assert_array_equal(levels[1][1], O.pyrdown_f32(dep0))
    gx, gy = O.sobel3(levels[2][0])
    np.testing.assert_array_equal(levels[2][2], gx)
    np.testing.assert_array_equal(levels[2][3], gy)
