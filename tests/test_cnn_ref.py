"""The numpy reference of the preconditioner network, pinned on CPU: its
convolution against scipy's correlate2d, its GELU against closed-form values,
its BF16 rounding against known bit patterns, and its loss gradient against
central finite differences."""
import numpy as np
import pytest
from scipy.signal import correlate2d

import cnn_ref as R


def test_conv_matches_scipy_correlate2d():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 7, 9))
    W = rng.normal(size=(2, 3, 3, 3))
    b = rng.normal(size=2)
    out = R.conv(x, W, b)
    for o in range(2):
        ref = b[o] + sum(correlate2d(x[c], W[o, c], mode="same") for c in range(3))
        np.testing.assert_allclose(out[o], ref, rtol=1e-12, atol=1e-12)


def test_conv_t_is_the_adjoint_of_conv():
    rng = np.random.default_rng(1)
    x, d = rng.normal(size=(4, 6, 5)), rng.normal(size=(3, 6, 5))
    W = rng.normal(size=(3, 4, 3, 3))
    lhs = np.sum(R.conv(x, W, np.zeros(3)) * d)
    rhs = np.sum(x * R.conv_t(d, W))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_gelu_values():
    # GELU(x) = x Phi(x): Phi(1) = 0.841344746..., Phi(-1) = 0.158655253...
    np.testing.assert_allclose(R.gelu(np.array([0.0, 1.0, -1.0, 3.0])),
                               [0.0, 0.8413447460685429, -0.15865525393145707,
                                3 * 0.9986501019683699], rtol=1e-12)
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    np.testing.assert_allclose(R.gelu_grad(z), (R.gelu(z + h) - R.gelu(z - h)) / (2 * h),
                               rtol=1e-7, atol=1e-9)


def test_bf16_rounding():
    # 1 + 2^-8 is halfway between BF16 1.0 and 1 + 2^-7: ties to even -> 1.0;
    # 1 + 3 * 2^-8 -> 1 + 2^-6 (even); 1 + 2^-8 + 2^-20 rounds up.
    v = R.to_bf16(np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5]))
    np.testing.assert_array_equal(v, [1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5])


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    H, W = 6, 7
    p = R.init_like_gpu_shapes(3)
    x = rng.normal(size=(4, H, W))
    t = rng.normal(size=(3, H, W))
    m = (rng.uniform(size=(H, W)) > 0.3).astype(float)
    L, g = R.loss_and_grad(p, x, t, m)
    idx = rng.choice(p.size, 40, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (R.loss_and_grad(pp, x, t, m)[0] - R.loss_and_grad(pm, x, t, m)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i
    assert R.num_params() == p.size == 20547
