"""Pinning of the U-Net oracle (CPU): BF16 rounding known answers and the
backward pass against central differences in float64."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_oracle as U  # noqa: E402


def test_bf16_rounding_known_answers():
    # BF16 keeps 8 significant bits: spacing 2^-7 on [1, 2); ties go to even
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159265, 65504.0, 2 ** -130],
                 np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -3.140625, 65536.0, 2 ** -130], np.float32)
    assert np.array_equal(U.bf16(x), want)


def test_parameter_count():
    # 9 * sum(cin*cout) of the 14 convolutions + 16 output weights
    assert U.n_params() == 9 * (2 * 16 + 16 * 16 + 16 * 32 + 32 * 32 + 32 * 64 + 64 * 64 + 64 * 128 + 128 * 128 +
                                192 * 64 + 64 * 64 + 96 * 32 + 32 * 32 + 48 * 16 + 16 * 16) + 16


def test_backward_matches_finite_differences():
    assert U.finite_difference_check(H=8, W=16, B=2, n_probe=16, seed=5) < 1e-4


def test_conv_is_a_correlation_with_zero_padding():
    # a single tap weight shifts the image; the border reads zeros
    x = np.arange(64, dtype=np.float32).reshape(1, 8, 8, 1)
    w = np.zeros((1, 9, 1), np.float32)
    w[0, 5, 0] = 1.0            # tap (dy=0, dx=+1)
    y = U.conv3x3(x, w)
    assert np.array_equal(y[0, :, :7, 0], x[0, :, 1:, 0])
    assert np.all(y[0, :, 7, 0] == 0)
