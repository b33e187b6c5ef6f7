"""The precision workload at full size and the benchmark fields against the oracle.

- 64^3 / Lmax 6 with the benchmark's smoothed fields (seed 1, std 5) against the oracle at
  reduced chain length (blocks 4/4/4), so the oracle's dense G^-1 tables fit.
- Lmax 8, lambda 0.5 on a grid with nz = 128 (the z solve of the 128^3 workload) against the
  oracle at small Ns.
- 128^3 / Lmax 8 / Ns 256 / lambda 0.5 (automatic hybrid store: lowest segments whole, the
  rest checkpointed): FP32 half spectrum within its stated bounds of FP64; w = 0 gives Q = 1
  and phi = f; equal block fields give the reversal identity q*(0) = P q(Ns); the hybrid
  store is bit-identical to every segment checkpointed.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _sweep(wlc, shape, box, Nblk, L, lam, w, **kw):
    g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, **kw)
    g.set_w(w)
    g.MDE_q()
    g.MDE_qstar()
    phi, Q, td = g.density()
    Ns = sum(Nblk)
    qN = g.get_q(+1, Ns)
    qs0 = g.get_q(-1, 0)
    g.close()
    return phi, Q, td, qN, qs0


def _oracle(oracle, shape, box, Nblk, L, lam, w):
    orc = oracle.Oracle(shape, box, Nblk, L, lam)
    q = orc.forward(w)
    qs = orc.backward(w)
    phi, Q, td = orc.density(q, qs)
    Ns = sum(Nblk)
    out = (phi, Q, td, orc.slice(q, Ns).copy(), orc.slice(qs, 0).copy())
    orc.close()
    return out


def test_benchmark_fields_match_oracle_at_reduced_chain_length(wlc, oracle):
    import workloads
    shape, box, Nblk, L, lam = (64, 64, 64), (5.0,) * 3, (4, 4, 4), (6, 6, 6), (1.0,) * 3
    w = workloads.smoothed_fields(shape)
    print("oracle: 64^3 Lmax 6, blocks 4/4/4, smoothed fields std 5", flush=True)
    ref = _oracle(oracle, shape, box, Nblk, L, lam, w)
    got = _sweep(wlc, shape, box, Nblk, L, lam, w)
    assert abs(got[1] - ref[1]) / ref[1] < 1e-10      # Q
    assert abs(got[2] - ref[2]) / ref[2] < 1e-10      # totden
    assert _rel(got[0], ref[0]) < 1e-10               # phi
    assert _rel(got[3], ref[3]) < 1e-10               # q(Ns)
    assert _rel(got[4], ref[4]) < 1e-10               # q*(0)


def test_lmax8_nz128_matches_oracle(wlc, oracle):
    shape, box, Nblk, L, lam = (128, 8, 16), (2.0, 1.0, 8.0), (3, 3, 3), (8, 8, 8), (0.5,) * 3
    rng = np.random.default_rng(12)
    w = 2.0 * rng.standard_normal((3,) + shape)
    print("oracle: 16x8x128 Lmax 8 lambda 0.5, blocks 3/3/3", flush=True)
    ref = _oracle(oracle, shape, box, Nblk, L, lam, w)
    got = _sweep(wlc, shape, box, Nblk, L, lam, w)
    assert abs(got[1] - ref[1]) / ref[1] < 1e-10
    assert _rel(got[0], ref[0]) < 1e-10
    assert _rel(got[3], ref[3]) < 1e-10
    assert _rel(got[4], ref[4]) < 1e-10


C3 = dict(shape=(128, 128, 128), box=(8.0,) * 3, L=(8, 8, 8), lam=(0.5,) * 3)


def _c3_blocks(wlc):
    return wlc.block_lengths(256, [1 / 3] * 3)


def test_precision_workload_fp32_spectrum_within_bounds(wlc):
    import workloads
    w = workloads.gyroid_fields(128)
    Nblk = _c3_blocks(wlc)
    print("128^3 Lmax 8 Ns 256: FP64 then FP32 half spectrum", flush=True)
    p64, Q64, _, _, _ = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"], w)
    p32, Q32, _, _, _ = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"], w,
                               spectrum_fp32=True)
    assert np.isfinite(Q64) and Q64 > 0
    assert _rel(p32, p64) < 1e-3          # measured 5.9e-4
    assert abs(Q32 - Q64) / Q64 < 5e-5    # measured 1.5e-5


def test_precision_workload_free_chain_and_reversal(wlc):
    import workloads
    Nblk = _c3_blocks(wlc)
    print("128^3 Lmax 8 Ns 256: w = 0, then equal block fields", flush=True)
    phi, Q, _, _, _ = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"],
                             np.zeros((3,) + C3["shape"]))
    assert abs(Q - 1.0) < 1e-12
    f = np.array(Nblk) / 256
    for b in range(3):
        assert np.abs(phi[b] - f[b]).max() < 1e-11
    base = workloads.gyroid_fields(128)[0]
    _, _, _, qN, qs0 = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"],
                              np.stack([base, base, base]))
    par = np.array([(-1) ** l for l in range(9) for m in range(-l, l + 1)])[:, None, None, None]
    assert _rel(qs0, par * qN) < 1e-10


def test_precision_workload_hybrid_store_equals_checkpoints(wlc, monkeypatch):
    import workloads
    w = workloads.gyroid_fields(128)
    Nblk = _c3_blocks(wlc)
    print("128^3 Lmax 8 Ns 256: hybrid store, then every segment checkpointed", flush=True)
    g = wlc.WormlikeMDE(C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"])
    stored = g.forward_stored()
    g.close()
    assert 0 < stored < 257   # the full store (350 GB) does not fit: a hybrid layout
    a = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"], w)
    monkeypatch.setenv("WLC_STORED_SEGMENTS", "0")
    b = _sweep(wlc, C3["shape"], C3["box"], Nblk, C3["L"], C3["lam"], w, ckpt_every=16)
    assert a[1] == b[1]
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[3], b[3])
    assert np.array_equal(a[4], b[4])
