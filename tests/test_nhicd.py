"""The bright-field NHICD list helpers (no GPU): the update-magnitude filter against a
2-D convolution with the same 5 x 5 window (scipy, zero boundary), which is the
outer product of the normalised 5-point Hamming window to the table's four digits,
and the list threshold against a full sort."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from mbirtem import capi

WINDOW = np.array([[0.0013, 0.0086, 0.0159, 0.0086, 0.0013],
                   [0.0086, 0.0581, 0.1076, 0.0581, 0.0086],
                   [0.0159, 0.1076, 0.1993, 0.1076, 0.0159],
                   [0.0086, 0.0581, 0.1076, 0.0581, 0.0086],
                   [0.0013, 0.0086, 0.0159, 0.0086, 0.0013]])


def test_window_is_a_separable_hamming_window():
    h = np.hamming(5)  # 0.08, 0.54, 1, 0.54, 0.08
    h = h / h.sum()
    np.testing.assert_allclose(np.outer(h, h), WINDOW, atol=6e-5)


def test_filter_is_a_zero_padded_convolution():
    rng = np.random.default_rng(1)
    mag = rng.exponential(1.0, (13, 17)).astype(np.float32)
    got = capi.nhicd_filter(mag)
    want = convolve2d(mag.astype(np.float64), WINDOW, mode="same", boundary="fill")
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    # a single spike spreads the window around it
    spike = np.zeros((9, 9), np.float32)
    spike[4, 4] = 2.0
    np.testing.assert_allclose(capi.nhicd_filter(spike)[2:7, 2:7], 2 * WINDOW, rtol=1e-12)


@pytest.mark.parametrize("n", [20, 37, 1000, 4096])
def test_threshold_is_the_95th_percentile_order_statistic(n):
    rng = np.random.default_rng(n)
    v = rng.normal(size=n)
    t = capi.nhicd_threshold(v)
    assert t == np.sort(v)[n - n // 20]
    # with distinct values the list (values above the threshold) holds n // 20 - 1 lines
    assert np.count_nonzero(v > t) == n // 20 - 1


def test_filter_of_a_large_map():
    """A map of more than 2^16 lines (filtered in row blocks on several host threads)
    against the same convolution."""
    rng = np.random.default_rng(2)
    mag = rng.exponential(1.0, (300, 260)).astype(np.float32)
    mag[rng.uniform(size=mag.shape) < 0.5] = 0.0
    got = capi.nhicd_filter(mag)
    want = convolve2d(mag.astype(np.float64), WINDOW, mode="same", boundary="fill")
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("kind", ["ties", "zeros", "signs", "tiny"])
def test_threshold_of_a_large_list(kind):
    """131072 values (a 512 x 256 map) with repeated values, many zeros, both signs
    and values near the smallest doubles: the order statistic of a full sort."""
    rng = np.random.default_rng(3)
    n = 131072
    v = {"ties": np.floor(rng.uniform(0, 8, n)),
         "zeros": np.where(rng.uniform(size=n) < 0.97, 0.0, rng.exponential(1.0, n)),
         "signs": rng.normal(size=n) * 10.0 ** rng.integers(-30, 30, n),
         "tiny": rng.uniform(size=n) * 1e-300}[kind]
    assert capi.nhicd_threshold(v) == np.sort(v)[n - n // 20]
