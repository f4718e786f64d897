"""Pins the FP64 restatement of the warm-start network (oracle/net.py):
the convolution against scipy.ndimage.correlate, its adjoint and weight
gradient by dot-product identities, the whole backward pass against central
finite differences, Adam's first step by its closed form, and the bfloat16
rounding by exact values."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import net


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 6, 7, 1))
    w = rng.normal(size=(27, 1, 1))
    k3 = w[:, 0, 0].reshape(3, 3, 3)
    want = ndimage.correlate(x[..., 0], k3, mode="constant", cval=0.0)
    np.testing.assert_allclose(net.conv(x, w)[..., 0], want, rtol=1e-12, atol=1e-12)


def test_conv_adjoint_and_weight_gradient_identities():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 5, 6, 3))
    w = rng.normal(size=(27, 4, 3))
    g = rng.normal(size=(4, 5, 6, 4))
    lhs = np.sum(net.conv(x, w) * g)
    assert abs(lhs - np.sum(x * net.conv_t(g, w))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(w * net.conv_wgrad(x, g))) < 1e-9 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = net.init_params(3)
    # make every layer matter and every relu partly active
    for p in params:
        p[0] *= 3.0
        p[1][...] = rng.normal(0.0, 0.05, p[1].shape)
    x = rng.uniform(0.2, 1.5, size=(4, 5, 6, 2))
    t = rng.uniform(0.2, 1.5, size=(4, 5, 6))
    loss, grads, _ = net.loss_and_grads(params, x, t)
    h = 1e-6
    for l in range(net.LAYERS):
        for i in range(2):
            flat = params[l][i].reshape(-1)
            for j in rng.choice(flat.size, size=min(4, flat.size), replace=False):
                old = flat[j]
                flat[j] = old + h
                lp = net.loss_and_grads(params, x, t)[0]
                flat[j] = old - h
                lm = net.loss_and_grads(params, x, t)[0]
                flat[j] = old
                fd = (lp - lm) / (2 * h)
                an = grads[l][i].reshape(-1)[j]
                assert abs(fd - an) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (l, i, j, fd, an)


def test_adam_first_step_closed_form():
    p = [[np.array([1.0, -2.0, 3.0]), np.array([0.5])]]
    g = [[np.array([0.3, -4.0, 0.0]), np.array([-1e-3])]]
    m = [[np.zeros(3), np.zeros(1)]]
    v = [[np.zeros(3), np.zeros(1)]]
    net.adam_step(p, g, m, v, 1, lr=0.01)
    # step 1: m_hat = g, v_hat = g^2 -> p -= lr * g / (|g| + eps)
    np.testing.assert_allclose(p[0][0], [1.0 - 0.01, -2.0 + 0.01, 3.0], rtol=0, atol=1e-7)
    np.testing.assert_allclose(p[0][1], [0.5 + 0.01], rtol=0, atol=1e-7)


def test_bf16_rounding_exact_values():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1e-40, 65504.0])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 0.0, 65536.0])
    got = net.bf16(vals)
    np.testing.assert_array_equal(got[:4], want[:4])
    assert abs(got[4] - 1e-40) < 2e-41  # subnormal: nearest bf16 is 2^-133
    assert got[5] == 65536.0
