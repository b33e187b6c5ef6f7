"""CPU oracle of the warm-start network's layers (test infrastructure only: the product never
imports it).

numpy restatement, in float64, of what the HIP training and inference kernels compute:
3x3x3 convolution with zero padding 1 and stride 1 (channels last, weights
[Cout][kz][ky][kx][Cin], bias per Cout), its gradients with respect to the input, the weights
and the bias, 2x2x2 average pooling and nearest 2x upsampling with their adjoints, ReLU,
and the mean-squared-error loss.  Operands that the device feeds to the bf16 matrix cores
are rounded to bfloat16 first (round to nearest even), exactly as the kernels round them;
the accumulation here is exact float64, the device's float32.

Pinned by: the adjoint identities <conv(x), y> = <x, conv^T(y)> and <pool(x), y> =
<x, pool^T(y)> (a transposed operand or a wrong flip breaks them), finite differences of
the loss (a dropped term or a wrong sign breaks them), and a direct quadruple-loop
convolution on tiny shapes.
"""
import numpy as np


def bf16(a):
    """Round float values to bfloat16 (round to nearest, ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def _pad(x):
    # x: [B, D, H, W, C] -> zero padded by 1 on D, H, W
    return np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))


def _patches(x):
    """[B, D, H, W, 27, C]: the 27 neighbours (kz, ky, kx order) of every voxel."""
    B, D, H, W, C = x.shape
    xp = _pad(x)
    out = np.empty((B, D, H, W, 27, C), dtype=x.dtype)
    t = 0
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out[:, :, :, :, t, :] = xp[:, kz:kz + D, ky:ky + H, kx:kx + W, :]
                t += 1
    return out


def conv3d(x, w, b):
    """y[v, co] = b[co] + sum_{tap, ci} w[co, tap, ci] x[v + tap - 1, ci]; w: [Cout, 27, Cin]."""
    P = _patches(x)
    return np.einsum("bdhwtc,otc->bdhwo", P, w, optimize=True) + b


def conv3d_backward(x, w, dy):
    """(dx, dw, db) of conv3d for the output gradient dy."""
    B, D, H, W, Cin = x.shape
    P = _patches(x)
    dw = np.einsum("bdhwo,bdhwtc->otc", dy, P, optimize=True)
    db = dy.sum(axis=(0, 1, 2, 3))
    # dx: correlation of dy with the flipped kernel, channels swapped
    wf = w[:, ::-1, :]  # tap t -> 26 - t flips all three axes
    Pd = _patches(dy)
    dx = np.einsum("bdhwto,otc->bdhwc", Pd, wf, optimize=True)
    return dx, dw, db


def pool2(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool2_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_backward(dy):
    B, D, H, W, C = dy.shape
    return dy.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def relu(x):
    return np.maximum(x, 0.0)


def mse(y, t):
    """loss = mean (y - t)^2 and its gradient."""
    d = y - t
    return float((d * d).mean()), 2.0 * d / d.size


def network_gradients(params, layers, x, t):
    """Loss and flat parameter gradients of the warm-start U-Net (the device's architecture
    and rounding points; layers: dicts with cin, cout, woff, boff, relu as the library
    reports them).  x [B, n, n, n, 11], t [B, n, n, n, 3]."""
    P = np.asarray(params, dtype=np.float64)

    def Wb(i):
        l = layers[i]
        w = P[l["woff"]:l["woff"] + l["cout"] * 27 * l["cin"]].reshape(l["cout"], 27, l["cin"])
        b = P[l["boff"]:l["boff"] + l["cout"]]
        return bf16(w), b

    B = x.shape[0]
    x0 = np.zeros(x.shape[:4] + (16,))
    x0[..., :11] = x
    x0 = bf16(x0)
    y = [None] * 11
    X = [None] * 11

    def conv(i, inp):
        w, b = Wb(i)
        X[i] = inp
        y[i] = conv3d(inp, w, b)
        return y[i]

    a0a = bf16(relu(conv(0, x0)))
    s0 = bf16(relu(conv(1, a0a)))
    p1 = bf16(pool2(s0))
    a1a = bf16(relu(conv(2, p1)))
    s1 = bf16(relu(conv(3, a1a)))
    p2 = bf16(pool2(s1))
    a2a = bf16(relu(conv(4, p2)))
    s2 = bf16(relu(conv(5, a2a)))
    p3 = bf16(pool2(s2))
    bt = bf16(relu(conv(6, p3)))
    c2 = np.concatenate([up2(bt), s2], axis=-1)
    d2 = bf16(relu(conv(7, c2)))
    c1 = np.concatenate([up2(d2), s1], axis=-1)
    d1 = bf16(relu(conv(8, c1)))
    c0 = np.concatenate([up2(d1), s0], axis=-1)
    d0 = bf16(relu(conv(9, c0)))
    out = conv(10, d0)
    d = out[..., :3] - t
    loss = float((d * d).mean())
    G = np.zeros_like(P)
    dy10 = np.zeros_like(out)
    dy10[..., :3] = 2.0 * d / d.size

    def back(i, dy, need_dx=True):
        l = layers[i]
        w, _ = Wb(i)
        dyb = bf16(dy)
        _, dw, _ = conv3d_backward(X[i], w, dyb)
        G[l["woff"]:l["woff"] + dw.size] = dw.reshape(-1)
        G[l["boff"]:l["boff"] + l["cout"]] = dy.sum(axis=(0, 1, 2, 3))
        if need_dx:
            dx, _, _ = conv3d_backward(X[i], w, dyb)
            return dx
        return None

    def mask(i, g):
        return g * (y[i] > 0) if layers[i]["relu"] else g

    dd0 = back(10, dy10)
    dc0 = back(9, mask(9, dd0))
    dd1 = up2_backward(dc0[..., :32])
    gs0 = dc0[..., 32:].copy()
    dc1 = back(8, mask(8, dd1))
    dd2 = up2_backward(dc1[..., :64])
    gs1 = dc1[..., 64:].copy()
    dc2 = back(7, mask(7, dd2))
    dbt = up2_backward(dc2[..., :128])
    gs2 = dc2[..., 128:].copy()
    dp3 = back(6, mask(6, dbt))
    gs2 += pool2_backward(dp3)
    da2a = back(5, mask(5, gs2))
    dp2 = back(4, mask(4, da2a))
    gs1 += pool2_backward(dp2)
    da1a = back(3, mask(3, gs1))
    dp1 = back(2, mask(2, da1a))
    gs0 += pool2_backward(dp1)
    da0a = back(1, mask(1, gs0))
    back(0, mask(0, da0a), need_dx=False)
    return loss, G, out[..., :3]
