"""Seeded synthetic fields of the benchmark and precision workloads (numpy PCG64, array order
(block, z, y, x), x fastest)."""
import numpy as np


def smoothed_fields(shape=(64, 64, 64), seed=1, sigma=2.0, std=5.0, nblk=3):
    """The 64^3 kernel-benchmark fields: i.i.d. N(0,1) per block, periodic Gaussian filter
    exp(-sigma^2 |2 pi m / n|^2 / 2) (sigma in cells), mean removed, each block scaled to the
    given standard deviation."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((nblk,) + tuple(shape))
    nz, ny, nx = shape
    kz = 2 * np.pi * np.fft.fftfreq(nz)[:, None, None]
    ky = 2 * np.pi * np.fft.fftfreq(ny)[None, :, None]
    kx = 2 * np.pi * np.fft.rfftfreq(nx)[None, None, :]
    filt = np.exp(-0.5 * sigma ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    for b in range(nblk):
        f = np.fft.irfftn(np.fft.rfftn(w[b]) * filt, s=shape, axes=(0, 1, 2))
        f -= f.mean()
        w[b] = f * (std / f.std())
    return w


def gyroid_fields(n=128, seed=2):
    """The 128^3 precision-workload fields: w_b = s_b (10 (sin X cos Y + sin Y cos Z +
    sin Z cos X) + 0.5 N(0,1)), X = 2 pi x / box, one N(0,1) field from the seed, block signs
    s = (+1, 0, -1)."""
    x = np.arange(n) * (2 * np.pi / n)
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    base = 10.0 * (np.sin(X) * np.cos(Y) + np.sin(Y) * np.cos(Z) + np.sin(Z) * np.cos(X))
    f = base + 0.5 * np.random.default_rng(seed).standard_normal((n, n, n))
    return np.stack([s * f for s in (1.0, 0.0, -1.0)])
