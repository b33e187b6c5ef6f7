"""Pins of the warm-start network's CPU oracle (oracle/unet_oracle.py): a direct loop
convolution, adjoint identities, finite differences of the loss, bf16 rounding."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_oracle as U  # noqa: E402


def _direct_conv(x, w, b):
    B, D, H, W, Cin = x.shape
    Cout = w.shape[0]
    y = np.zeros((B, D, H, W, Cout)) + b
    for bb in range(B):
        for d in range(D):
            for h in range(H):
                for ww in range(W):
                    t = 0
                    for kz in range(3):
                        for ky in range(3):
                            for kx in range(3):
                                zz, yy, xx = d + kz - 1, h + ky - 1, ww + kx - 1
                                if 0 <= zz < D and 0 <= yy < H and 0 <= xx < W:
                                    y[bb, d, h, ww] += w[:, t, :] @ x[bb, zz, yy, xx]
                                t += 1
    return y


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 5, 3))
    w = rng.standard_normal((4, 27, 3))
    b = rng.standard_normal(4)
    assert np.abs(U.conv3d(x, w, b) - _direct_conv(x, w, b)).max() < 1e-12


def test_conv_adjoint_identities():
    """<conv(x), y> = <x, dx(y)> and <conv(x), y> = <w, dw(x, y)> (bias zero): the input and
    weight gradients are the exact adjoints of the linear maps."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 4, 6, 5))
    w = rng.standard_normal((3, 27, 5))
    y = rng.standard_normal((2, 4, 4, 6, 3))
    lhs = (U.conv3d(x, w, np.zeros(3)) * y).sum()
    dx, dw, db = U.conv3d_backward(x, w, y)
    assert lhs == pytest.approx((x * dx).sum(), rel=1e-12)
    assert lhs == pytest.approx((w * dw).sum(), rel=1e-12)
    assert np.allclose(db, y.sum(axis=(0, 1, 2, 3)))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 6, 8, 2))
    y = rng.standard_normal((1, 2, 3, 4, 2))
    assert (U.pool2(x) * y).sum() == pytest.approx((x * U.pool2_backward(y)).sum(), rel=1e-12)
    z = rng.standard_normal((1, 4, 6, 8, 2))
    assert (U.up2(y) * z).sum() == pytest.approx((y * U.up2_backward(z)).sum(), rel=1e-12)
    assert np.allclose(U.pool2(U.up2(y)), y)


def test_loss_gradient_by_finite_differences():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 3, 3, 4, 2))
    w = rng.standard_normal((2, 27, 2)) * 0.3
    b = rng.standard_normal(2)
    t = rng.standard_normal((1, 3, 3, 4, 2))

    def loss(w_):
        return U.mse(U.relu(U.conv3d(x, w_, b)), t)[0]

    y = U.conv3d(x, w, b)
    _, g = U.mse(U.relu(y), t)
    g = g * (y > 0)
    _, dw, _ = U.conv3d_backward(x, w, g)
    h = 1e-6
    for idx in [(0, 0, 0), (1, 13, 1), (0, 26, 1), (1, 5, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        assert (loss(wp) - loss(wm)) / (2 * h) == pytest.approx(dw[idx], rel=1e-5, abs=1e-9)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0])
    r = U.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0 + 2 ** -6            # tie -> even (up)
    assert r[4] == -2.5 and r[5] == 0.0
