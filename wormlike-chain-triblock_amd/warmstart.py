"""Learned warm start for batched SCFT: data, training and evaluation.

The network (wlc_unet_*, hand-written bf16 matrix-core kernels) maps an instance's seed fields
and parameters to proposed initial fields; plain SCFT from them still decides the converged state.

  features(case)      11 input channels: the 3 seed fields x 10 (unit variance for the
                      0.1 N(0,1) seeds), then f_A, f_B, f_C, chiN_AB / 30, chiN_BC / 30,
                      chiN_AC / 30, lambda / 4, box / 6 broadcast over the grid
  target(w)           converged fields with each block's mean removed (a constant per block
                      is a gauge of the melt: every chain carries the same block fractions,
                      so it rescales Q and leaves phi and F unchanged), / 10
  proposal(y)         fields = 10 y, used as the initial w of SCFT

generate(...) runs the batching workload's generator at small grids through the batched solver
and keeps the final fields of the instances that converged or came close; train(...) fits the
network; predict(...) proposes initial fields for new instances at any grid edge that is a
multiple of 8 (the network is fully convolutional).
"""
import time

import numpy as np

import batch_case as B

SCALE = 10.0


def draw_at(n_inst, seed, edges, Ls, Ns=B.NS):
    """The batching workload's parameter draws with the grid edge and Lmax drawn from the given
    sets (training data at small grids)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_inst):
        edge = int(edges[rng.integers(len(edges))])
        L = int(Ls[rng.integers(len(Ls))])
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


def solve(W, cases, init=None, max_iter=520, tol=1e-4, graph=True, log=print):
    """Batched SCFT of cases from their seed fields (or init[i]): 20 simple-mixing iterations
    (0.05), then Anderson over 10 histories mixing the inputs (0.1); per-instance results."""
    b = W.Batch(cases, W.scft_opts(max_iter=max_iter, n_simple=20, n_hist=10, lam_simple=0.05,
                                   lam_anderson=0.1, form=1, tol=tol))
    for i, c in enumerate(cases):
        b.set_w(i, c["w"] if init is None else init[i])
    t0 = time.perf_counter()
    it = 0
    while True:
        act = b.iterate(20, use_graph=graph)
        it += 20
        log(f"solve: {it} iterations, {act} of {len(cases)} advancing, {time.perf_counter() - t0:.1f} s")
        if act == 0:
            break
    wall = time.perf_counter() - t0
    res = []
    for i in range(len(cases)):
        st = b.status(i)
        st["w"] = b.get_w(i)
        res.append(st)
    b.close()
    return res, wall


def generate(W, n_inst, seed, edges, Ls, chunk=64, max_iter=520, keep_below=1e-3, log=print):
    """(features, target) pairs from batched SCFT of draw_at(...) instances: the final fields of
    every instance whose largest residual fell below keep_below (converged to 1e-4 or close to
    it after max_iter iterations), grouped by grid edge."""
    data = {}
    info = []
    cases = draw_at(n_inst, seed, edges, Ls)
    for k in range(0, n_inst, chunk):
        part = cases[k:k + chunk]
        res, wall = solve(W, part, max_iter=max_iter, log=log)
        kept = 0
        for c, r in zip(part, res):
            ok = np.isfinite(r["F"]) and max(r["resid"]) < keep_below
            if ok:
                X, T = data.setdefault(c["shape"][0], ([], []))
                X.append(features(c))
                T.append(target(r["w"]))
                kept += 1
            info.append(dict(edge=c["shape"][0], L=c["L"][0], iterations=r["iterations"],
                             converged=r["converged"], resid=max(r["resid"]), kept=bool(ok)))
        log(f"generate: instances {k}..{k + len(part) - 1}: {kept} kept, {wall:.1f} s")
    return {e: (np.stack(X), np.stack(T)) for e, (X, T) in data.items()}, info


def train(W, data, epochs, batch=4, lr=1e-3, seed=0, log=print, net=None):
    """Adam with a cosine-decayed rate over mini-batches of one grid size at a time (the network
    is fully convolutional; it is resized between sizes).  data: {edge: (X, T)}."""
    first = min(data)
    if net is None:
        net = W.UNet(first, batch, seed=seed)
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    nsamp = sum(len(X) for X, _ in data.values())
    for ep in range(epochs):
        rate = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        batches = []
        for e in rng.permutation(sorted(data)):  # sizes in random order, each size's batches together
            X = data[int(e)][0]
            order = rng.permutation(len(X))
            batches += [(int(e), order[k:k + batch]) for k in range(0, len(X) - batch + 1, batch)]
        tot = 0.0
        for e, idx in batches:
            if net.n != e or net.B != batch:
                net.resize(e, batch)
            X, T = data[e]
            tot += net.train_step(X[idx], T[idx], rate)
        hist.append(tot / max(1, len(batches)))
        if ep % 5 == 0 or ep == epochs - 1:
            log(f"train: epoch {ep} loss {hist[-1]:.5f} ({nsamp} samples, {time.perf_counter() - t0:.1f} s)")
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
