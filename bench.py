"""Benchmark: MDE contour-steps/s of the wormlike-chain propagator solve on one GPU.

Workload (the kernel benchmark configuration of the reference's MDE path): 64^3
periodic grid, box 5.0, three blocks of equal fraction (Ns = 200 -> 67/67/66 contour
steps), Lmax = 6 (49 harmonics) and lambda = 1.0 in every block; w_A,B,C are i.i.d.
N(0,1) from seed 1, smoothed by a periodic Gaussian filter of width 2 cells and
scaled to standard deviation 5; q(0) = delta_{l0}.  One bench step is one forward
sweep, one backward sweep (with the fused contour quadrature) and the density
reduction: 2 * Ns = 400 contour steps.  Inputs are resident in HBM before the timed
region starts.

Prints one JSON line; see the repository's design notes for the field definitions.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "wormlike-chain-triblock_amd"))

METRIC = "MDE contour-steps/s at 64^3 Lmax=6"
PEAK_HBM_GBS = 8000.0
KERNELS = ["rhs_fft_xy", "zsolve", "ifft_xy", "density_acc", "inv_rhs_fft_xy"]


def workload_w(shape=(64, 64, 64), seed=1, sigma=2.0, std=5.0, nblk=3):
    """Seeded block fields: N(0,1), periodic Gaussian filter exp(-sigma^2 |2 pi m/n|^2 / 2)
    in cells, then each block scaled to the given standard deviation."""
    import workloads
    return workloads.smoothed_fields(shape, seed, sigma, std, nblk)


def algorithmic_bytes(M, N, Nk):
    """FP64 bytes each kernel must move per launch (BDF3 step); the step total is
    the 4 M N 8 + 8 N + 64 M Nk figure of the field traffic."""
    return {
        "rhs_fft_xy": 3 * M * N * 8 + N * 8 + M * Nk * 16,
        "zsolve": 2 * M * Nk * 16,
        "ifft_xy": M * Nk * 16 + M * N * 8,
        "density_acc": 2 * M * N * 8 + 2 * N * 8,
        # inverse xy of step s fused with the rhs + forward xy of step s+1: q(s) stays on
        # chip, so only two history slices are read
        "inv_rhs_fft_xy": M * Nk * 16 + M * N * 8 + 2 * M * N * 8 + N * 8 + M * Nk * 16,
        "step": 4 * M * N * 8 + 8 * N + 64 * M * Nk,
    }


def read_traffic(path):
    """Per-launch HBM bytes per kernel from a committed rocprofv3 summary: lines
    'traffic <kernel> <bytes>'."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 3 and p[0] == "traffic":
                out[p[1]] = float(p[2])
    return out


def cpu_baseline(w, target_s=15.0):
    """The oracle (C restatement of the reference's per-step dataflow: BDF3 RHS,
    per-harmonic R2C FFT, dense G^-1 apply via its lower triangle, C2R FFT) timed on
    this host on a bounded sample of the same workload: BDF3 steps of block A."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle  # noqa: E402  (checker / CPU baseline only)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc = oracle.Oracle((64, 64, 64), (5.0, 5.0, 5.0), (67, 67, 66), (6, 6, 6), (1.0, 1.0, 1.0),
                        nthreads=threads)
    t0 = time.time()
    first = orc.time_bdf3_steps(w[0], 1)  # builds the G^-1 table (excluded) + 1 step
    build_s = time.time() - t0 - first
    n = max(2, min(64, int(target_s / max(first, 1e-3))))
    dt = orc.time_bdf3_steps(w[0], n)
    # the same steps on one core (a few steps: ~1/3 of the budget)
    oracle.Oracle.set_threads(1)
    one = orc.time_bdf3_steps(w[0], 1)
    n1 = max(1, min(16, int(target_s / 3 / max(one, 1e-3))))
    dt1 = orc.time_bdf3_steps(w[0], n1)
    oracle.Oracle.set_threads(threads)
    orc.close()
    return {"value": n / dt, "unit": "contour-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} BDF3 contour steps of the 64^3 Lmax=6 workload (block A field), "
                      f"dense G^-1 table build of {build_s:.1f} s excluded",
            "single_core": {"value": n1 / dt1, "sample": f"{n1} of the same steps on 1 core"}}


def workload_w128(n=128, seed=2):
    """w_b = s_b (10 (sin X cos Y + sin Y cos Z + sin Z cos X) + 0.5 N(0,1)), X = 2 pi x / box,
    one N(0,1) field from the seed, block signs s = (+1, 0, -1)."""
    import workloads
    return workloads.gyroid_fields(n, seed)


def bench_grid128(W, args):
    """FP64 forward + backward sweeps of the 128^3 Lmax=8 workload (349 GB of forward store
    per sweep would not fit: checkpoints every 16 contour steps, segments recomputed during
    the backward sweep).  With --spectrum-fp32 the same sweeps also run with the half
    spectrum stored as FP32, and the line reports that rate with its phi / Q / F (chiN = 20)
    differences from the FP64 run."""
    n, Ns, L, lam, box, ck = 128, 256, 8, 0.5, 8.0, 16
    Nblk = W.block_lengths(Ns, [1 / 3, 1 / 3, 1 / 3])
    M, N, Nk = (L + 1) ** 2, n ** 3, n * n * (n // 2 + 1)
    w = workload_w128(n)
    chi = np.full((3, 3), 20.0) - 20.0 * np.eye(3)
    steps = max(1, min(args.steps, 3))
    runs = {}
    for fp32 in ((False, True) if args.spectrum_fp32 else (False,)):
        # automatic forward store: 16-step segments, the lowest ones kept whole as far as
        # device memory allows (the full 350 GB store does not fit)
        g = W.WormlikeMDE((n,) * 3, (box,) * 3, Nblk, (L,) * 3, (lam,) * 3, spectrum_fp32=fp32)
        g.set_w(w)
        stored = g.forward_stored()
        for _ in range(min(args.warmup, 1)):
            g.MDE_q()
            g.MDE_qstar()
            g.density_device()
        g.sync()
        g.event_record(0)
        Q = td = None
        for _ in range(steps):
            g.MDE_q()
            g.MDE_qstar()
            Q, td = g.density_device()
        g.event_record(1)
        ms = g.event_elapsed(0, 1)
        g.sync()
        if not (np.isfinite(Q) and np.isfinite(td)):
            raise RuntimeError("non-finite Q/totden")
        g.set_timing(2)
        g.MDE_q()
        g.MDE_qstar()
        phi, Q, _ = g.density()
        kms, kcnt = g.get_timing()
        g.set_timing(0)
        g.close()
        runs[fp32] = (ms, kms, kcnt, phi, Q, free_energy(w, phi, Q, chi, box), stored)
    fp32 = args.spectrum_fp32
    ms, kms, kcnt, phi, Q, F, stored = runs[fp32]
    contour_steps = 2 * Ns * steps
    step_s = ms / 1e3 / contour_steps
    b_step = algorithmic_bytes(M, N, Nk)["step"] - (32 * M * Nk if fp32 else 0)
    out = {
        "metric": "MDE contour-steps/s at 128^3 Lmax=8" + (", FP32 half spectrum" if fp32 else ""),
        "value": round(contour_steps / (ms / 1e3), 2),
        "unit": "contour-steps/s", "n_gpus": 1, "steps": steps, "ms_per_step": round(ms / steps, 2),
        "higher_is_better": True,
        "dtype": "f64 arithmetic, f32 spectrum storage" if fp32 else "f64", "data": "synthetic",
        "config": {"workload": "MDE forward+backward sweep, 128^3, box 8.0, Lmax=8 (81 harmonics), "
                               "Ns=256, lambda=0.5, w = s_b (10 (sin X cos Y + sin Y cos Z + sin Z cos X) "
                               "+ 0.5 N(0,1) seed 2), s = (+,0,-); forward store in 16-step segments, the "
                               "lowest kept whole as far as device memory allows, the rest "
                               "checkpointed and recomputed; one step = one sweep of 512 contour steps",
                   "grid": [n, n, n], "Lmax": L, "Ns": Ns, "blocks": Nblk, "segment": ck,
                   "forward_slices_stored": stored},
        "step_roofline": {"algorithmic_bytes": b_step, "achieved": round(b_step / step_s / 1e9, 1),
                          "frac": round(b_step / step_s / 1e9 / PEAK_HBM_GBS, 4)},
        "kernels_avg_us": {k: round(1e3 * m / c, 1) if c else 0.0 for k, m, c in zip(KERNELS, kms, kcnt)},
        "Q": Q,
    }
    if fp32:
        _, _, _, phi64, Q64, F64, _ = runs[False]
        out["fp64_value"] = round(contour_steps / (runs[False][0] / 1e3), 2)
        out["error_vs_fp64"] = {"phi_max_rel": float(np.abs(phi - phi64).max() / np.abs(phi64).max()),
                                "Q_rel": abs(Q - Q64) / Q64, "F_abs_kT": abs(F - F64),
                                "F_fp64": F64, "chiN": 20.0}
    print(json.dumps(out))


def free_energy(w, phi, Q, chi, box):
    """F = -ln Q + (1/V) int [sum_{a<b} chi_ab phi_a phi_b - sum_a w_a phi_a - xi (1 - sum phi)],
    xi = mean of the block fields, with the periodic-Simpson volume weights (the ABC form
    of SCFT_Density_FE.f90:296-298,333)."""
    nz, ny, nx = phi.shape[1:]
    c = [np.where(np.arange(n) % 2 == 0, 2.0 / 3.0, 4.0 / 3.0) for n in (nz, ny, nx)]
    wgt = c[0][:, None, None] * c[1][None, :, None] * c[2][None, None, :] / (nx * ny * nz)
    nb = len(phi)
    xi = w.mean(0)
    e = -xi * (1.0 - phi.sum(0))
    for a in range(nb):
        e -= w[a] * phi[a]
        for b in range(a + 1, nb):
            e += chi[a][b] * phi[a] * phi[b]
    return -np.log(Q) + float((wgt * e).sum())


def bench_spectrum_fp32(W, args):
    """The 64^3 Lmax=6 sweep with the half spectrum stored as FP32: contour-steps/s, and
    the differences of phi, Q and F (chiN = 20 for every pair) from an FP64 run."""
    shape, Ns, L, lam, box = (64, 64, 64), 200, 6, 1.0, 5.0
    Nblk = W.block_lengths(Ns, [1 / 3, 1 / 3, 1 / 3])
    w = workload_w(shape)
    chi = np.full((3, 3), 20.0) - 20.0 * np.eye(3)
    out = {}
    for fp32 in (False, True):
        g = W.WormlikeMDE(shape, (box,) * 3, Nblk, (L,) * 3, (lam,) * 3, spectrum_fp32=fp32)
        g.set_w(w)
        for _ in range(args.warmup):
            g.MDE_q()
            g.MDE_qstar()
            g.density_device()
        g.sync()
        g.event_record(0)
        for _ in range(args.steps):
            g.MDE_q()
            g.MDE_qstar()
            g.density_device()
        g.event_record(1)
        ms = g.event_elapsed(0, 1)
        g.MDE_q()
        g.MDE_qstar()
        phi, Q, _ = g.density()
        g.close()
        out[fp32] = (2 * Ns * args.steps / (ms / 1e3), phi, Q, free_energy(w, phi, Q, chi, box))
    v64, phi64, Q64, F64 = out[False]
    v32, phi32, Q32, F32 = out[True]
    step_bytes = 32 * 49 * 64 ** 3 + 8 * 64 ** 3 + 32 * 49 * 64 * 64 * 33
    print(json.dumps({
        "metric": "MDE contour-steps/s at 64^3 Lmax=6, FP32 half spectrum", "value": round(v32, 1),
        "unit": "contour-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "higher_is_better": True, "dtype": "f64 arithmetic, f32 spectrum storage",
        "data": "synthetic", "fp64_value": round(v64, 1),
        "config": {"workload": "the 64^3 benchmark sweep (box 5.0, Lmax=6, Ns=200, lambda=1.0, "
                               "smoothed N(0,1) w seed 1, std 5); the half spectrum between the "
                               "FFT passes and the solve stored as FP32", "grid": [64, 64, 64],
                   "Lmax": L, "Ns": Ns},
        "step_roofline": {"algorithmic_bytes": step_bytes,
                          "frac": round(step_bytes * v32 / 1e9 / PEAK_HBM_GBS, 4)},
        "error_vs_fp64": {"phi_max_rel": float(np.abs(phi32 - phi64).max() / np.abs(phi64).max()),
                          "Q_rel": abs(Q32 - Q64) / Q64, "F_abs_kT": abs(F32 - F64), "F_fp64": F64,
                          "chiN": 20.0},
    }))


# field update of both time-to-residual runs (plain and batched): 20 simple-mixing iterations
# (step 0.05), then Anderson over 10 histories mixing the inputs with step 0.1 (the reference's
# form, mixing the outputs, diverges on most of these instances)
MIX = dict(n_simple=20, n_hist=10, lam_s=0.05, lam_a=0.1, form=1)


def c4_cases(args):
    import batch_case as B
    lo, hi = (int(x) for x in args.c4_range.split(":")) if args.c4_range else (0, args.c4)
    out = list(enumerate(B.draw()))[lo:hi]
    if args.c4_edges:
        keep = {int(e) for e in args.c4_edges.split(",")}
        out = [(i, c) for i, c in out if c["shape"][0] in keep]
    return out


def c4_report(kind, rows, extra):
    """JSON line of a time-to-residual run over the batching workload's instances."""
    t = np.array([r["time_s"] for r in rows])
    it = np.array([r["iterations"] for r in rows])
    conv = [r["converged"] for r in rows]
    tc = t[np.array(conv, dtype=bool)] if any(conv) else np.array([np.nan])
    out = {
        "metric": f"time-to-residual {extra['tol']:g}, {kind}",
        "value": round(float(np.mean(t)), 4), "unit": "s", "higher_is_better": False,
        "n_gpus": 1, "dtype": "f64", "data": "synthetic",
        "instances": len(rows), "converged": int(sum(conv)),
        "mean_s": round(float(np.mean(t)), 4), "median_s": round(float(np.median(t)), 4),
        "mean_s_converged": round(float(np.mean(tc)), 4),
        "mean_iterations": round(float(np.mean(it)), 1), "median_iterations": float(np.median(it)),
        "config": {"workload": "linear ABC wormlike triblocks from seed 3 (the 48-instance batching "
                               "draw): cubic grid edge {24,32,48,64}, Lmax {4,6}, f_A, f_B ~ "
                               "U[0.2,0.45], chiN_ij ~ U[12,30], lambda ~ U[0.25,4], box ~ U[3,6], "
                               "Ns 100, w = 0.1 N(0,1); n_simple simple-mixing iterations (step "
                               "lam_s) then Anderson (n_hist histories, form 1: inputs mixed, step "
                               "lam_a), at most max_iter iterations", "tol": extra.pop("tol"),
                   "max_iter": extra.pop("max_iter"),
                   "mixing": extra.pop("mixing", dict(MIX))},
        "per_instance": [{k: (round(v, 10) if isinstance(v, float) else v) for k, v in r.items()}
                         for r in rows],
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


def bench_plain_c4(args):
    """Each instance alone through the Fortran driver (module wormlike_MDE on one handle,
    field update on the host), one after the other; time = the driver's loop wall time."""
    import scft_case as S
    rows = []
    for i, c in c4_cases(args):
        r, info = S.run_case(method=1, max_iter=args.max_iter, tol=args.tol, case=c, timeout=900, **MIX)
        rows.append({"i": i, "edge": c["shape"][0], "L": c["L"][0], "iterations": info["iterations"],
                     "converged": bool(info["converged"]), "time_s": info["seconds"],
                     "F": float(r[-1, 1]) if len(r) else None, "error": info["error"]})
        print(f"# plain {i}: {rows[-1]}", file=sys.stderr, flush=True)
    c4_report("plain single-instance SCFT (Fortran driver, one instance at a time)", rows,
              {"tol": args.tol, "max_iter": args.max_iter})


def unet_flops(layers, n, B):
    """Algorithmic FLOPs of one forward pass of the warm-start network (real channels, not the
    16-channel padding): 2 * voxels * 27 * Cin * Cout per convolution."""
    tot = 0
    for l in layers:
        vox = B * (n >> l["level"]) ** 3
        tot += 2 * vox * 27 * l["cin_real"] * l["cout_real"]
    return tot


def bench_unet(args, W):
    """Training and inference throughput of the warm-start network at edge 64 (one instance
    per batch, the size of the largest batching instances): wall time per train step (forward,
    loss, backward, Adam) and per forward pass, host copies of inputs and outputs included;
    FLOPs counted on the real channels (a training step: forward, input gradient except for
    the first layer, weight gradient)."""
    n, B = args.unet_edge, 1
    net = W.UNet(n, B, seed=0)
    layers = net.layers()
    rng = np.random.default_rng(0)
    x = rng.standard_normal((B, n, n, n, 11)).astype(np.float32)
    t = 0.1 * rng.standard_normal((B, n, n, n, 3)).astype(np.float32)
    for _ in range(args.warmup):
        net.train_step(x, t, 1e-4)
        net.forward(x)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(x, t, 1e-4)
    t_train = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.forward(x)
    t_fwd = (time.perf_counter() - t0) / args.steps
    f_fwd = unet_flops(layers, n, B)
    l0 = layers[0]
    f_train = 3 * f_fwd - 2 * B * n ** 3 * 27 * l0["cin_real"] * l0["cout_real"]
    params = net.num_params()
    net.close()
    print(json.dumps({
        "metric": "warm-start network throughput", "unit": "TFLOP/s", "dtype": "bf16 (fp32 accumulate)",
        "config": {"workload": f"U-Net, edge {n}, batch {B}, 11 input / 3 output channels, "
                               f"{len(layers)} 3x3x3 convolutions", "params": params},
        "forward_ms": round(1e3 * t_fwd, 3), "train_step_ms": round(1e3 * t_train, 3),
        "forward_tflops": round(f_fwd / t_fwd / 1e12, 3), "train_tflops": round(f_train / t_train / 1e12, 3),
        "forward_flop": f_fwd, "train_flop": f_train, "peak_tflops_bf16_dense": 2500.0,
        "steps": args.steps, "warmup": args.warmup}), flush=True)


def bench_warm_train(args, W):
    """Training data from the batched solver (draws of the batching workload's generator at
    grid edges 24 and 32, Lmax 4 or 6, seed 11), then training of the warm-start network on
    them; writes the weights to --out."""
    import warmstart as WS
    log = lambda m: print("# " + m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    data, info = WS.generate(W, args.warm_instances, 11, (24, 32), (4, 6), chunk=64, log=log)
    t_gen = time.perf_counter() - t0
    net, hist, t_train = WS.train(W, data, args.warm_epochs, batch=4, lr=1e-3, seed=0, log=log)
    np.save(args.out, net.params())
    print(json.dumps({"metric": "warm-start network training", "samples": {int(e): len(X) for e, (X, _) in data.items()},
                      "instances": len(info), "generate_s": round(t_gen, 1), "train_s": round(t_train, 1),
                      "epochs": args.warm_epochs, "loss": [round(float(h), 6) for h in hist],
                      "params": net.num_params(), "weights": args.out}), flush=True)
    net.close()


def bench_batch_c4(args, W):
    """All instances in one batch (wlc_batch: one launch per kernel variant and contour step
    for every instance, field update on the device, one SCFT iteration replayed from a HIP
    graph); time of an instance = wall time from the first iteration to the end of the
    iteration in which it converged (or stopped)."""
    idx_cases = c4_cases(args)
    cases = [c for _, c in idx_cases]
    mix = dict(MIX)
    if args.warm:
        mix["n_simple"] = 0  # the proposal is close to the solution: Anderson from the start
    if args.n_simple is not None:
        mix["n_simple"] = args.n_simple
    if args.lam_anderson is not None:
        mix["lam_a"] = args.lam_anderson
    b = W.Batch(cases, W.scft_opts(max_iter=args.max_iter, n_simple=mix["n_simple"], n_hist=mix["n_hist"],
                                   lam_simple=mix["lam_s"], lam_anderson=mix["lam_a"],
                                   form=mix["form"], tol=args.tol))
    t_net = 0.0
    init = [c["w"] for c in cases]
    if args.warm:
        import warmstart as WS
        net = W.UNet(cases[0]["shape"][0], 1)
        net.set_params(np.load(args.warm))
        t1 = time.perf_counter()
        init = WS.predict(net, cases)
        t_net = time.perf_counter() - t1
        net.close()
    for i, c in enumerate(cases):
        b.set_w(i, init[i])
    print(f"# batch of {len(cases)} instances created, {b.launches(True)} launches per iteration"
          + (f", warm start from {args.warm} ({t_net:.2f} s)" if args.warm else ""),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    done = 0
    while True:
        active = b.iterate(10, use_graph=not args.no_graph)
        done += 10
        print(f"# {done} iterations, {active} instances advancing, {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if active == 0:
            break
    wall = time.perf_counter() - t0
    rows = []
    for k, (i, c) in enumerate(idx_cases):
        st = b.status(k)
        rows.append({"i": i, "edge": c["shape"][0], "L": c["L"][0], "iterations": st["iterations"],
                     "converged": st["converged"], "time_s": st["time_ms"] / 1e3 + t_net,
                     "F": st["F"], "resid": max(st["resid"])})
    launches = b.launches(True)
    b.close()
    kind = ("batched SCFT (all instances per launch, device field update, "
            + ("direct launches)" if args.no_graph else "HIP graph)"))
    if args.warm:
        kind = "batched warm-started SCFT (network proposal as initial fields; " + kind[14:]
    c4_report(kind, rows,
              {"batch_wall_s": round(wall + t_net, 3),
               "amortised_s_per_instance": round((wall + t_net) / len(rows), 4),
               "network_s": round(t_net, 3), "launches_per_iteration": launches, "tol": args.tol,
               "max_iter": args.max_iter, "mixing": mix})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed sweeps (each 2*Ns contour steps)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--traffic-file", default=os.path.join(HERE, "profiles", "traffic_64cube_L6.txt"))
    ap.add_argument("--scft", action="store_true",
                    help="instead of the sweep benchmark: plain SCFT of the 32^3 ABC case "
                         "(Lmax 4, Ns 100) through the Fortran driver -- 50 simple-mixing "
                         "iterations reported per iteration, then a simple-then-Anderson run "
                         "towards residual 1e-6 (value = its time if it converges)")
    ap.add_argument("--grid128", action="store_true",
                    help="instead of the 64^3 benchmark: the FP64 sweep of the 128^3 Lmax=8 "
                         "workload (box 8, Ns 256, lambda 0.5, forward store in 16-step segments, "
                         "the lowest kept whole as far as device memory allows), one JSON line "
                         "of its own")
    ap.add_argument("--spectrum-fp32", action="store_true",
                    help="the 64^3 sweep with the half spectrum stored as FP32 (all arithmetic "
                         "FP64), one JSON line of its own with the phi / Q / F differences "
                         "from an FP64 run on the same fields")
    ap.add_argument("--plain-c4", action="store_true",
                    help="time-to-residual of the batching workload's instances run one at a "
                         "time through the Fortran driver (the plain comparator)")
    ap.add_argument("--batch-c4", action="store_true",
                    help="time-to-residual of the same instances advanced together in one batch")
    ap.add_argument("--no-graph", action="store_true",
                    help="--batch-c4: launch every kernel of an iteration directly instead of "
                         "replaying the captured HIP graph (same kernels and order; for kernel "
                         "tracers)")
    ap.add_argument("--c4", type=int, default=48, help="instances of the batching workload")
    ap.add_argument("--c4-range", default="", help="lo:hi slice of the batching workload's instances")
    ap.add_argument("--tol", type=float, default=1e-6, help="residual of the time-to-residual runs")
    ap.add_argument("--max-iter", type=int, default=520, help="iteration cap of the time-to-residual runs")
    ap.add_argument("--c4-edges", default="", help="only instances with these grid edges (e.g. 24,32)")
    ap.add_argument("--warm", default="", help="weights (.npy) of the warm-start network for --batch-c4")
    ap.add_argument("--n-simple", type=int, default=None,
                    help="--batch-c4: simple-mixing iterations before Anderson (default 20, 0 with --warm)")
    ap.add_argument("--lam-anderson", type=float, default=None,
                    help="--batch-c4: Anderson step (default 0.1, inputs mixed)")
    ap.add_argument("--warm-train", action="store_true",
                    help="generate training data with the batched solver and train the warm-start "
                         "network; weights to --out")
    ap.add_argument("--warm-instances", type=int, default=256)
    ap.add_argument("--warm-epochs", type=int, default=30)
    ap.add_argument("--out", default="unet_weights.npy")
    ap.add_argument("--unet", action="store_true",
                    help="instead of the sweep benchmark: training and inference throughput of the "
                         "warm-start network at --unet-edge (one JSON line)")
    ap.add_argument("--unet-edge", type=int, default=64)
    ap.add_argument("--copy-only", action="store_true",
                    help="only run three 256 MiB device-to-device copies (the counter "
                         "calibration kernel of wormlike-chain-triblock_amd/profile_summary.py "
                         "traffic; run as a process of its own)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the propagator solve runs on a single device "
              f"(multi-GPU sharding is a design note, not built); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.scft:
        import scft_case as S
        rows, info0 = S.run_case(method=0, max_iter=50, lam_s=0.05)
        if info0["error"]:
            raise RuntimeError(info0["error"])
        rows1, info1 = S.run_case(method=1, max_iter=1500, n_simple=MIX["n_simple"], n_hist=MIX["n_hist"],
                                  lam_s=MIX["lam_s"], lam_a=MIX["lam_a"], tol=1e-6, form=MIX["form"])
        done = len(rows1)
        print(json.dumps({
            "metric": "time-to-residual 1e-6, plain SCFT (32^3 ABC wormlike triblock, Lmax=4, Ns=100)",
            "value": round(info1["seconds"], 3) if info1["converged"] else None, "unit": "s",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
            "converged": info1["converged"], "iterations": done,
            "final_residual": float(rows1[-1, 3]) if done else None,
            "final_F": float(rows1[-1, 1]) if done else None,
            "error": info1["error"],
            "seconds_per_iteration": round(info0["seconds"] / 50, 5),
            "config": {"workload": "32^3, box 4.0, f=1/3 each, chiN=20 all pairs, lambda=1, "
                                   "Lmax=4, Ns=100 (33/33/34), w=0.1 N(0,1) seed 0 +/-2cos(2 pi z/box) "
                                   "on A/C; 20 simple-mixing iterations (step 0.05) then Anderson "
                                   "(10 histories, inputs mixed, lambda 0.1), at most 1500 iterations",
                       "mixing": "simple then Anderson"},
            "simple_mixing_50": {"F": rows[:, 1].round(10).tolist(),
                                 "lnQ": rows[:, 2].round(10).tolist(),
                                 "phi_mean": rows[:, 7:10].round(10).tolist(),
                                 "residual": rows[:, 3].round(10).tolist(),
                                 "seconds": round(info0["seconds"], 3)},
        }))
        return

    if args.copy_only:
        import torch
        a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 3}))
        return

    if args.plain_c4:
        bench_plain_c4(args)
        return

    import wlc_mde as W

    if args.batch_c4:
        bench_batch_c4(args, W)
        return
    if args.warm_train:
        bench_warm_train(args, W)
        return
    if args.unet:
        bench_unet(args, W)
        return
    if args.grid128:
        bench_grid128(W, args)
        return
    if args.spectrum_fp32:
        bench_spectrum_fp32(W, args)
        return

    shape = (64, 64, 64)
    Ns = 200
    Nblk = W.block_lengths(Ns, [1 / 3, 1 / 3, 1 / 3])
    L, lam, box = 6, 1.0, 5.0
    M = (L + 1) ** 2
    N = 64 ** 3
    Nk = 64 * 64 * 33
    w = workload_w(shape)
    g = W.WormlikeMDE(shape, (box,) * 3, Nblk, (L,) * 3, (lam,) * 3)
    g.set_w(w)  # host -> HBM before the timed region

    for _ in range(args.warmup):
        g.MDE_q()
        g.MDE_qstar()
        g.density_device()
    g.sync()
    t0 = time.perf_counter()
    g.event_record(0)
    Q = td = None
    for _ in range(args.steps):
        g.MDE_q()
        g.MDE_qstar()
        Q, td = g.density_device()
    g.event_record(1)
    ms = g.event_elapsed(0, 1)
    g.sync()
    wall = time.perf_counter() - t0
    if not (np.isfinite(Q) and np.isfinite(td)):
        raise RuntimeError("non-finite Q/totden")
    contour_steps = 2 * Ns * args.steps
    value = contour_steps / (ms / 1e3)

    # Per-kernel launch durations: one more sweep pair with HIP events around every launch
    # on the stream it runs on.  In the timed sweeps above the forward and backward sweeps
    # run concurrently, so a kernel there shares the GPU with the other sweep's kernels;
    # here (timing mode 2) the backward sweep waits for the forward one, so each step
    # kernel has the GPU to itself and bytes / duration is that kernel's own rate.
    g.set_timing(2)
    g.MDE_q()
    g.MDE_qstar()
    g.density_device()
    kms, kcnt = g.get_timing()
    g.set_timing(0)
    avg_us = {k: (1e3 * m / c if c else 0.0) for k, m, c in zip(KERNELS, kms, kcnt)}
    ab = algorithmic_bytes(M, N, Nk)
    dom = max(KERNELS, key=lambda k: kms[KERNELS.index(k)])
    achieved = ab[dom] / (avg_us[dom] * 1e-6) / 1e9
    traffic = read_traffic(args.traffic_file).get(dom)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel": dom,
            "algorithmic_bytes": ab[dom], "avg_launch_us": round(avg_us[dom], 2)}
    step_s = ms / 1e3 / contour_steps
    step_roof = {"algorithmic_bytes": ab["step"], "achieved": round(ab["step"] / step_s / 1e9, 1),
                 "frac": round(ab["step"] / step_s / 1e9 / PEAK_HBM_GBS, 4)}
    g.close()

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "contour-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "MDE forward+backward sweep, 64^3, box 5.0, Lmax=6 (49 harmonics), "
                               "Ns=200 (67/67/66), lambda=1.0, smoothed N(0,1) w (seed 1, sigma 2 "
                               "cells, std 5); one step = one sweep of 400 contour steps",
                   "grid": [64, 64, 64], "Lmax": L, "Ns": Ns, "blocks": Nblk,
                   "parallelism": "1 GPU"},
        "roofline": roof,
        "step_roofline": step_roof,
        "kernels_avg_us": {k: round(v, 2) for k, v in avg_us.items()},
        "wall_s": round(wall, 3),
        "Q": Q,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
