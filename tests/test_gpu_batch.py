"""Batched instances (wlc_batch_*): one launch per kernel variant and contour step covers
every instance.  Checked against per-instance handles (the single-instance path, itself
oracle-checked) on mixed grids, Lmax, block structures and chain lengths; the on-device
field update against the Fortran host driver (module scft_iterate), which the CPU oracle
loop pins; graph replays against eager launches."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "wormlike-chain-triblock_amd", "fortran", "scft_driver")

CHI3 = np.array([[0.0, 20.0, 15.0], [20.0, 0.0, 25.0], [15.0, 25.0, 0.0]])

# mixed instances: radix-3 edges, 64 x 64 planes with nz = 64 (whole-tile z solve at Lmax 6),
# the 32 x 32 specialised planes, Lmax growing and shrinking at junctions, 1-3 blocks,
# uneven chain lengths
CASES = [
    dict(shape=(16, 16, 16), box=(3.0, 3.0, 3.0), Nblk=(4, 5, 3), L=(3, 2, 3), lam=(1.0, 0.5, 2.0)),
    dict(shape=(24, 24, 24), box=(4.0, 4.0, 4.0), Nblk=(7, 6, 7), L=(4, 4, 4), lam=(0.8,) * 3),
    dict(shape=(12, 12, 12), box=(2.5, 2.5, 2.5), Nblk=(9, 8), L=(2, 5), lam=(1.5, 0.3)),
    dict(shape=(32, 32, 32), box=(4.5, 4.5, 4.5), Nblk=(6, 7, 7), L=(6, 6, 6), lam=(1.0,) * 3),
    dict(shape=(64, 64, 64), box=(5.0, 5.0, 5.0), Nblk=(4, 3, 5), L=(6, 6, 6), lam=(1.0,) * 3),
    dict(shape=(8, 12, 16), box=(2.0, 3.0, 4.0), Nblk=(11,), L=(4,), lam=(2.0,)),
    dict(shape=(48, 48, 48), box=(5.5, 5.5, 5.5), Nblk=(5, 5, 4), L=(4, 4, 4), lam=(0.4,) * 3),
    dict(shape=(24, 24, 24), box=(3.3, 3.3, 3.3), Nblk=(8, 5, 7), L=(4, 4, 4), lam=(3.0,) * 3),
    dict(shape=(16, 16, 32), box=(2.0, 2.0, 4.0), Nblk=(3, 4, 6), L=(1, 3, 2), lam=(1.0, 1.0, 0.5)),
]


def _chi(nb):
    return CHI3[:nb, :nb] if nb < 3 else CHI3


def _w(c, seed):
    rng = np.random.default_rng(seed)
    return 1.5 * rng.standard_normal((len(c["Nblk"]),) + c["shape"])


def test_batch_sweep_matches_single_instance_handles(wlc, monkeypatch):
    """Every forward slice, Q, totden and phi against a handle on the same instance, with the
    handle's sweeps sequential (the batch accumulates the density terms in the sequential
    order).  The batched twin of a kernel runs the same source on arguments read from the
    job table; the compiler may contract multiply-adds differently in the two, so the results
    agree to rounding (observed <= 3.4e-15 relative), not always bit for bit."""
    monkeypatch.setenv("WLC_CONCURRENT", "0")
    cases = [dict(c, chi=np.pad(_chi(len(c["Nblk"])), ((0, 3 - len(c["Nblk"])),) * 2)) for c in CASES]
    print(f"batch of {len(cases)} instances", flush=True)
    b = wlc.Batch(cases)
    ws = [_w(c, 10 + i) for i, c in enumerate(cases)]
    for i, w in enumerate(ws):
        b.set_w(i, w)
    b.sweep()
    for i, c in enumerate(cases):
        phi, Q, td = b.density(i)
        g = wlc.WormlikeMDE(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
        g.set_w(ws[i])
        g.MDE_q()
        g.MDE_qstar()
        gphi, gQ, gtd = g.density()
        Ns = sum(c["Nblk"])
        for s in sorted({1, 2, 3, c["Nblk"][0], Ns // 2, Ns}):
            M = g.slice_M(s)
            ref = g.get_q(+1, s)
            assert np.abs(b.get_q(i, s, M) - ref).max() <= 1e-13 * np.abs(ref).max(), (i, s)
        g.close()
        assert abs(Q - gQ) <= 1e-13 * gQ, i
        assert abs(td - gtd) <= 1e-13 * gtd, i
        assert np.abs(phi - gphi).max() <= 1e-13 * np.abs(gphi).max(), i
    b.close()


def _write_input(path, c, chi, w, max_iter, n_simple, n_hist, lam_s, lam_a, tol, form):
    nz, ny, nx = c["shape"]
    nb = len(c["Nblk"])
    hdr = np.zeros(16, dtype=np.int32)
    hdr[:5] = [nx, ny, nz, nb, sum(c["Nblk"])]
    hdr[5:5 + nb] = c["Nblk"]
    hdr[8:8 + nb] = c["L"]
    hdr[11:16] = [max_iter, 1, n_simple, n_hist, form]
    lam3 = np.zeros(3)
    lam3[:nb] = c["lam"]
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.asarray(c["box"], dtype=np.float64).tobytes())
        f.write(lam3.tobytes())
        f.write(np.asarray(chi, dtype=np.float64).reshape(9).tobytes())
        f.write(np.asarray([lam_s, lam_a, tol], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(w, dtype=np.float64).tobytes())


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (no flang)")
@pytest.mark.parametrize("form", [0, 1])
def test_batch_field_update_matches_the_fortran_driver(wlc, tmp_path, form):
    """Three instances, 2 simple-mixing then 6 Anderson iterations (history 3, the ring
    wraps): the device update follows the host module scft_iterate iteration by iteration.
    The two differ only in rounding (reduction order; the Anderson system from a Gram matrix
    instead of direct differences), which the mixing amplifies: the first iterations agree to
    1e-11 kT in F, all eight to 1e-7 (observed 3.4e-9), the final fields to 1e-6."""
    sel = [CASES[0], CASES[2], CASES[8]]
    kw = dict(max_iter=8, n_simple=2, n_hist=3, lam_s=0.05, lam_a=0.2, tol=1e-14, form=form)
    cases = []
    for c in sel:
        chi = np.zeros((3, 3))
        nb = len(c["Nblk"])
        chi[:nb, :nb] = _chi(nb)
        cases.append(dict(c, chi=chi))
    opts = wlc.scft_opts(max_iter=kw["max_iter"], n_simple=kw["n_simple"], n_hist=kw["n_hist"],
                         form=form, lam_simple=kw["lam_s"], lam_anderson=kw["lam_a"], tol=kw["tol"])
    b = wlc.Batch(cases, opts)
    ws = [0.5 * _w(c, 40 + i) for i, c in enumerate(cases)]
    for i, w in enumerate(ws):
        b.set_w(i, w)
    traj = [[] for _ in cases]
    for _ in range(kw["max_iter"]):
        b.iterate(1, use_graph=True)
        for i in range(len(cases)):
            traj[i].append(b.status(i))
    for i, c in enumerate(cases):
        st = traj[i][-1]
        assert st["iterations"] == kw["max_iter"] and st["done"] and not st["converged"]
        wb = b.get_w(i)
        d = tmp_path / f"i{i}"
        d.mkdir()
        _write_input(d / "in.bin", c, c["chi"], ws[i], **kw)
        r = subprocess.run([EXE, str(d / "in.bin"), str(d / "out.txt"), str(d / "w.bin")],
                           capture_output=True, text=True, timeout=120, cwd=d)
        assert r.returncode == 0, r.stdout + r.stderr
        rows = np.array([[float(x) for x in ln.split()] for ln in open(d / "out.txt")
                         if not ln.startswith("#")])
        wf = np.fromfile(d / "w.bin").reshape(ws[i].shape)
        nb = len(c["Nblk"])
        for k, (t, row) in enumerate(zip(traj[i], rows)):
            assert t["iterations"] == k + 1
            tol_F = 1e-11 if k < 3 else 1e-7
            assert abs(t["F"] - row[1]) < tol_F, (i, k)
            assert abs(t["lnQ"] - row[2]) < tol_F, (i, k)
            assert np.allclose(t["resid"][:nb], row[4:4 + nb], rtol=1e-9 if k < 3 else 1e-5, atol=0), (i, k)
        assert np.abs(wb - wf).max() < 1e-6 * np.abs(wf).max()
    b.close()


def test_graph_replays_match_eager_launches(wlc):
    """The same iterations replayed from the captured graph (several replays, host reads
    between them) and launched eagerly give bit-identical fields and states."""
    cases = []
    for c in (CASES[0], CASES[1], CASES[5]):
        chi = np.zeros((3, 3))
        nb = len(c["Nblk"])
        chi[:nb, :nb] = _chi(nb)
        cases.append(dict(c, chi=chi))
    opts = wlc.scft_opts(max_iter=40, n_simple=3, n_hist=4, tol=1e-14)
    out = []
    for use_graph in (False, True):
        b = wlc.Batch(cases, opts)
        for i, c in enumerate(cases):
            b.set_w(i, 0.5 * _w(c, 70 + i))
        for _ in range(4):
            b.iterate(2, use_graph=use_graph)
            _ = [b.status(i) for i in range(len(cases))]
        out.append(([b.get_w(i) for i in range(len(cases))],
                    [b.status(i) for i in range(len(cases))]))
        b.close()
    for i in range(len(cases)):
        assert np.array_equal(out[0][0][i], out[1][0][i])
        for k in ("iterations", "F", "lnQ", "resid"):
            assert out[0][1][i][k] == out[1][1][i][k]


def test_converged_instances_stop_advancing(wlc):
    """An instance whose residual falls below tol stops: its fields and iteration count stay
    fixed while the others go on (a uniform field at w = 0 is a fixed point of the update:
    residual 0 at the first iteration)."""
    cases = []
    for c in (CASES[0], CASES[5]):
        chi = np.zeros((3, 3))
        nb = len(c["Nblk"])
        chi[:nb, :nb] = _chi(nb)
        cases.append(dict(c, chi=chi))
    opts = wlc.scft_opts(max_iter=30, n_simple=2, n_hist=3, tol=1e-8)
    b = wlc.Batch(cases, opts)
    b.set_w(0, 0.5 * _w(cases[0], 5))
    b.set_w(1, np.zeros((1,) + cases[1]["shape"]))  # homopolymer at w = 0: t = 0 exactly
    b.iterate(3)
    s1 = b.status(1)
    assert s1["converged"] and s1["iterations"] == 1 and s1["time_ms"] > 0
    w1 = b.get_w(1)
    b.iterate(5)
    assert b.status(1)["iterations"] == 1
    assert np.array_equal(b.get_w(1), w1)
    assert b.status(0)["iterations"] == 8
    b.close()


def test_compacted_launch_sequence_matches_the_full_one(wlc):
    """Once at most half of the instances still advance, the batch rebuilds its launch
    sequence (and graph) over those alone; the advancing instances run the same kernels on
    the same data, so every field, iteration count and free energy equals a batch that keeps
    every instance in the sequence, bit for bit."""
    cases = []
    for c in (CASES[0], CASES[5], CASES[1], CASES[5], CASES[5]):
        chi = np.zeros((3, 3))
        nb = len(c["Nblk"])
        chi[:nb, :nb] = _chi(nb)
        cases.append(dict(c, chi=chi))
    out = []
    for compact in (True, False):
        b = wlc.Batch(cases, wlc.scft_opts(max_iter=12, n_simple=2, n_hist=3, tol=1e-8, compact=compact))
        for i, c in enumerate(cases):
            if i in (1, 3, 4):
                # homopolymers at w = 0 converge at the first iteration: 2 of 5 remain, so the
                # sequence is rebuilt over those two
                b.set_w(i, np.zeros((1,) + c["shape"]))
            else:
                b.set_w(i, 0.5 * _w(c, 5 + i))
        b.iterate(4, use_graph=True)
        b.iterate(10, use_graph=True)
        out.append([(b.get_w(i), b.status(i)) for i in range(len(cases))])
        b.close()
    for (w0, s0), (w1, s1) in zip(*out):
        assert np.array_equal(w0, w1)
        assert s0["iterations"] == s1["iterations"] and s0["converged"] == s1["converged"]
        assert s0["F"] == s1["F"] and s0["lnQ"] == s1["lnQ"]


@pytest.mark.parametrize("policy", ["2", "3"])
def test_two_lanes_match_one_lane(wlc, monkeypatch, policy):
    """The sweeps of the two lanes run on two streams at once; every instance still runs the
    same kernels on the same data in the same order, so fields, iteration counts, free
    energies and forward slices equal the single-stream sequence (WLC_BATCH_LANES=1) bit for
    bit, for whole-group and instance-by-instance lanes, eager and replayed from the graph."""
    cases = []
    for c in (CASES[0], CASES[1], CASES[5], CASES[6], CASES[2], CASES[7]):
        chi = np.zeros((3, 3))
        nb = len(c["Nblk"])
        chi[:nb, :nb] = _chi(nb)
        cases.append(dict(c, chi=chi))
    out = []
    for lanes in ("1", policy):
        monkeypatch.setenv("WLC_BATCH_LANES", lanes)
        b = wlc.Batch(cases, wlc.scft_opts(max_iter=20, n_simple=2, n_hist=3, tol=1e-14))
        for i, c in enumerate(cases):
            b.set_w(i, 0.5 * _w(c, 40 + i))
        b.iterate(2, use_graph=False)
        b.iterate(3, use_graph=True)
        # slice 2 lies in every instance's first block
        out.append([(b.get_w(i), b.status(i), b.get_q(i, 2, (c["L"][0] + 1) ** 2))
                    for i, c in enumerate(cases)])
        b.close()
    for (w0, s0, q0), (w1, s1, q1) in zip(*out):
        assert np.array_equal(w0, w1) and np.array_equal(q0, q1)
        assert s0["iterations"] == s1["iterations"]
        assert s0["F"] == s1["F"] and s0["lnQ"] == s1["lnQ"] and s0["resid"] == s1["resid"]
    # instances with a nonzero force advance all 5 iterations (a one-block homopolymer may
    # stop at its first)
    assert max(s["iterations"] for _, s, _ in out[0]) == 5
