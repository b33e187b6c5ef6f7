"""Learned warm start for batched SCFT: data, training and evaluation.

The network (wlc_unet_*, hand-written bf16 matrix-core kernels) maps an instance's seed fields
and parameters to proposed initial fields; plain SCFT from the seed fields stays the judge.

  features(case)      11 input channels: the 3 seed fields x 10 (unit variance for the
                      0.1 N(0,1) seeds), then f_A, f_B, f_C, chiN_AB / 30, chiN_BC / 30,
                      chiN_AC / 30, lambda / 4, box / 6 broadcast over the grid
  target(w)           converged fields with each block's mean removed (a constant per block
                      is a gauge of the melt: every chain carries the same block fractions,
                      so it rescales Q and leaves phi and F unchanged), / 10
  proposal(y)         fields = 10 y, used as the initial w of SCFT

generate(...) runs the batching workload's generator at a fixed grid and Lmax through the
batched solver and keeps the converged instances; train(...) fits the network; evaluate(...)
runs held-out draws from their seed fields and from the network's proposal through the same
batched solver and reports iterations, time and the free-energy difference.
"""
import time

import numpy as np

import batch_case as B

SCALE = 10.0


def draw_at(n_inst, seed, edge, L, Ns=B.NS):
    """The batching workload's parameter draws with the grid edge and Lmax fixed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        fA = float(rng.uniform(0.2, 0.45))
        fB = float(rng.uniform(0.2, 0.45))
        fC = 1.0 - fA - fB
        cab, cbc, cac = (float(v) for v in rng.uniform(12.0, 30.0, 3))
        lam = float(rng.uniform(0.25, 4.0))
        box = float(rng.uniform(3.0, 6.0))
        w = 0.1 * rng.standard_normal((3, edge, edge, edge))
        chi = np.array([[0.0, cab, cac], [cab, 0.0, cbc], [cac, cbc, 0.0]])
        out.append(dict(shape=(edge,) * 3, box=(box,) * 3, Nblk=B.block_lengths(Ns, (fA, fB, fC)),
                        L=(L,) * 3, lam=(lam,) * 3, chi=chi, f=(fA, fB, fC), w=w))
    return out


def features(c):
    n = c["shape"][0]
    x = np.empty((n, n, n, 11), dtype=np.float32)
    for b in range(3):
        x[..., b] = SCALE * c["w"][b]
    chi = c["chi"]
    s = [c["f"][0], c["f"][1], c["f"][2], chi[0, 1] / 30, chi[1, 2] / 30, chi[0, 2] / 30,
         c["lam"][0] / 4, c["box"][0] / 6]
    for k, v in enumerate(s):
        x[..., 3 + k] = v
    return x


def target(w):
    t = np.stack([(w[b] - w[b].mean()) / SCALE for b in range(3)], axis=-1)
    return t.astype(np.float32)


def proposal(y):
    """[n, n, n, 3] network output -> [3, n, n, n] fields."""
    return SCALE * np.moveaxis(np.asarray(y, dtype=np.float64), -1, 0)


def solve(W, cases, init=None, max_iter=520, graph=True):
    """Batched SCFT of cases from their seed fields (or init[i]); per-instance results."""
    b = W.Batch(cases, W.scft_opts(max_iter=max_iter, n_simple=20, n_hist=10, lam_simple=0.05,
                                   lam_anderson=0.1, tol=1e-6))
    for i, c in enumerate(cases):
        b.set_w(i, c["w"] if init is None else init[i])
    t0 = time.perf_counter()
    while b.iterate(20, use_graph=graph) > 0:
        pass
    wall = time.perf_counter() - t0
    res = []
    for i in range(len(cases)):
        st = b.status(i)
        st["w"] = b.get_w(i)
        res.append(st)
    b.close()
    return res, wall


def generate(W, n_inst, seed, edge, L, chunk=64, log=print):
    """Converged (features, target) pairs from batched SCFT of draw_at(...) instances."""
    X, T, info = [], [], []
    cases = draw_at(n_inst, seed, edge, L)
    for k in range(0, n_inst, chunk):
        part = cases[k:k + chunk]
        res, wall = solve(W, part)
        nconv = 0
        for c, r in zip(part, res):
            if r["converged"]:
                X.append(features(c))
                T.append(target(r["w"]))
                nconv += 1
            info.append(dict(iterations=r["iterations"], converged=r["converged"], F=r["F"]))
        log(f"generate: instances {k}..{k + len(part) - 1}: {nconv} converged, {wall:.1f} s")
    return np.stack(X), np.stack(T), info


def train(W, X, T, epochs, batch=4, lr=1e-3, seed=0, log=print, net=None):
    n = X.shape[1]
    if net is None:
        net = W.UNet(n, batch, seed=seed)
    else:
        net.resize(n, batch)
    rng = np.random.default_rng(seed)
    N = len(X)
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(N)
        rate = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        tot = 0.0
        for k in range(0, N - batch + 1, batch):
            idx = order[k:k + batch]
            tot += net.train_step(X[idx], T[idx], rate)
        hist.append(tot / max(1, N // batch))
        if ep % 10 == 0 or ep == epochs - 1:
            log(f"train: epoch {ep} loss {hist[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    return net, hist, time.perf_counter() - t0


def predict(net, cases):
    """Proposed initial fields for each case (one network call per instance)."""
    out = []
    for c in cases:
        n = c["shape"][0]
        if net.n != n or net.B != 1:
            net.resize(n, 1)
        y = net.forward(features(c)[None])
        out.append(proposal(y[0]))
    return out
