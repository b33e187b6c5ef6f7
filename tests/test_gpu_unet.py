"""The warm-start network's kernels against the CPU oracle (oracle/unet_oracle.py): one
convolution and its three gradients on bf16 matrix cores, then the whole network's loss,
output and parameter gradients, then that training lowers the loss."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_oracle as U  # noqa: E402


@pytest.mark.parametrize("shape,cin,cout", [((2, 4, 6, 8), 16, 16), ((1, 8, 8, 8), 32, 48),
                                            ((3, 5, 3, 7), 48, 16), ((1, 8, 16, 16), 64, 32),
                                            ((2, 4, 4, 16), 32, 128), ((1, 3, 8, 24), 16, 64)])
def test_conv_and_gradients_match_oracle(wlc, shape, cin, cout):
    """Shapes with an x edge that is a multiple of 8 (and 32 | voxels) take the transposed
    weight-gradient kernel, the others the element-wise one; input channels that are a multiple
    of 32 take the 16x16x32 matrix-core step, 16 and 48 the 16x16x16 one; 16, 32 and 64 or more
    output channels give 1, 2 and 4 output tiles per workgroup."""
    rng = np.random.default_rng(cin + cout)
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((cout, 27, cin)) * 0.2
    b = rng.standard_normal(cout)
    dy = rng.standard_normal(shape + (cout,))
    y, dx, dw, db = wlc.test_conv3d(x, w, b, dy)
    xb, wb, dyb = U.bf16(x), U.bf16(w), U.bf16(dy)
    ry = U.conv3d(xb, wb, b.astype(np.float32).astype(np.float64))
    rdx, rdw, _ = U.conv3d_backward(xb, wb, dyb)
    rdb = dy.astype(np.float32).astype(np.float64).sum(axis=(0, 1, 2, 3))
    for got, ref in ((y, ry), (dx, rdx), (dw, rdw), (db, rdb)):
        assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()


def test_lds_staged_convolution_equals_the_direct_one(wlc, monkeypatch):
    """The forward kernel with the weights staged through LDS (default for Cin a multiple of
    32) runs the same matrix-core steps in the same order as the direct kernel
    (WLC_CONV_LDS=0): outputs and input gradients are equal bit for bit."""
    rng = np.random.default_rng(4)
    shape, cin, cout = (1, 8, 16, 16), 64, 128
    x = rng.standard_normal(shape + (cin,))
    w = rng.standard_normal((cout, 27, cin)) * 0.2
    b = rng.standard_normal(cout)
    dy = rng.standard_normal(shape + (cout,))
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("WLC_CONV_LDS", flag)
        out.append(wlc.test_conv3d(x, w, b, dy))
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def _inputs(n, B, seed):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.standard_normal((B, n, n, n, 3)),
                        np.broadcast_to(rng.uniform(0, 1, (B, 1, 1, 1, 8)), (B, n, n, n, 8))], -1)
    t = rng.standard_normal((B, n, n, n, 3))
    return x.astype(np.float32), t.astype(np.float32)


def test_network_gradients_match_oracle(wlc):
    n, B = 8, 2
    net = wlc.UNet(n, B, seed=5)
    layers = net.layers()
    x, t = _inputs(n, B, 6)
    loss = net.gradients(x, t)
    g = net.grads()
    p = net.params()
    y = net.forward(x)
    net.close()
    rloss, rg, ry = U.network_gradients(p, layers, x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(rloss, rel=2e-3)
    assert np.abs(y - ry).max() <= 1e-2 * np.abs(ry).max()
    for i, l in enumerate(layers):
        nw = l["cout"] * 27 * l["cin"]
        for a, b in ((l["woff"], l["woff"] + nw), (l["boff"], l["boff"] + l["cout"])):
            ref = rg[a:b]
            err = np.linalg.norm(g[a:b] - ref) / max(np.linalg.norm(ref), 1e-30)
            assert err < 2e-2, (i, err)


def test_training_lowers_the_loss(wlc):
    n, B = 16, 2
    net = wlc.UNet(n, B, seed=1)
    assert 1.8e6 < net.num_params() < 2.0e6
    x, _ = _inputs(n, B, 3)
    t = 0.5 * x[..., :3] + 0.1  # a learnable map of the input
    l0 = net.train_step(x, t, 1e-3)
    for _ in range(100):
        l1 = net.train_step(x, t, 1e-3)
    net.close()
    assert np.isfinite(l1) and l1 < 0.3 * l0   # observed 0.30 after 60 steps
