"""The linear resize of the recorded fluid velocity (simulate.py:24-28,
jax.image.resize(..., method="linear")): the oracle's restatement against
closed forms, and the device's float32 taps (mc_resize_taps, host only)
against the oracle's float64 weights."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import step_ref as R


@pytest.mark.parametrize("n", [1, 5, 32])
def test_equal_sizes_are_the_identity(n):
    np.testing.assert_array_equal(R.resize_weights(n, n), np.eye(n))


@pytest.mark.parametrize("n_in,n_out", [(8, 16), (16, 8), (64, 48), (48, 64), (7, 3), (3, 11)])
def test_weights_are_a_partition_of_unity(n_in, n_out):
    w = R.resize_weights(n_in, n_out)
    np.testing.assert_allclose(w.sum(axis=0), 1.0, rtol=0, atol=1e-12)
    assert (w >= 0).all()


def test_upsampling_interpolates_a_ramp_exactly():
    n = 12
    f = 3.0 + 0.5 * np.arange(n)
    out = R.resize_linear(np.tile(f[None, :], (n, 1))[None], (1, n, 2 * n))[0, 0]
    o = np.arange(2 * n)
    sample = (o + 0.5) / 2.0 - 0.5
    want = 3.0 + 0.5 * np.clip(sample, 0, n - 1)
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-12)


def test_downsampling_by_two_is_the_1331_filter():
    w = R.resize_weights(16, 8)
    for o in range(1, 7):
        col = w[:, o]
        np.testing.assert_allclose(col[2 * o - 1:2 * o + 3], [1 / 8, 3 / 8, 3 / 8, 1 / 8],
                                   rtol=0, atol=1e-12)
        assert col.sum() - col[2 * o - 1:2 * o + 3].sum() == 0.0
    # the first output's window is cut by the border and renormalised
    np.testing.assert_allclose(w[:3, 0], np.array([3, 3, 1]) / 7, rtol=0, atol=1e-12)


def test_separable_resize_of_a_product_field():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=10), rng.normal(size=14)
    u = np.stack([np.outer(a, b), 2 * np.outer(a, b)])
    out = R.resize_linear(u, (2, 5, 21))
    ra = R.resize_weights(10, 5).T @ a
    rb = R.resize_weights(14, 21).T @ b
    np.testing.assert_allclose(out[0], np.outer(ra, rb), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[1], 2 * out[0], rtol=1e-12, atol=1e-12)


def _device_taps(n_in, n_out):
    from microcosmos_amd import _lib
    lib = _lib.load()
    off = np.zeros(n_out + 1, np.int32)
    cap = 4 * n_in + 4 * n_out + 8
    idx = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    n = lib.mc_resize_taps(n_in, n_out, off.ctypes.data, idx.ctypes.data, w.ctypes.data, cap)
    assert 0 <= n <= cap
    dense = np.zeros((n_in, n_out))
    for o in range(n_out):
        for k in range(off[o], off[o + 1]):
            dense[idx[k], o] += w[k]
    return dense


@pytest.mark.parametrize("n_in,n_out", [(8, 16), (16, 8), (64, 32), (32, 64), (100, 48),
                                         (48, 100), (1024, 512), (7, 3)])
def test_device_taps_match_the_oracle(n_in, n_out):
    """float32 like jax's own tabulation: equal to the float32 restatement,
    and within float32 rounding of the float64 weights."""
    d = _device_taps(n_in, n_out)
    w32 = R.resize_weights(n_in, n_out, np.float32)
    assert ((d != 0) == (w32 != 0)).all()
    np.testing.assert_array_equal(d.astype(np.float32), w32)
    w64 = R.resize_weights(n_in, n_out)
    np.testing.assert_allclose(d, w64, rtol=0, atol=5e-6)


def test_device_taps_reject_bad_sizes():
    from microcosmos_amd import _lib
    lib = _lib.load()
    off = np.zeros(4, np.int32)
    assert lib.mc_resize_taps(0, 3, off.ctypes.data, None, None, 0) == -1
    assert lib.mc_resize_taps(3, 0, off.ctypes.data, None, None, 0) == -1
