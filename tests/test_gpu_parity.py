"""GPU parity of the HIP propagator solve against the CPU oracle (through the C ABI).

Tolerance (FP64 path): 1e-10 relative in phi, Q and every propagator slice, the
parity bar stated for the FP64 GPU path.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-10
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("L", list(range(9)))
def test_k_aligned_solve_matches_dense_inverse(wlc, oracle, L):
    rng = np.random.default_rng(100 + L)
    M = (L + 1) ** 2
    k = rng.standard_normal((64, 3)) * 25
    k[0] = 0.0                       # k = 0: diagonal G
    k[1, :2] = 0.0                   # k along z: no rotation needed
    k[2, 2] = 0.0                    # k in the xy plane
    k[3] = [40.0, 0.0, 0.0]
    v = rng.standard_normal((64, M)) + 1j * rng.standard_normal((64, M))
    for cd, sigma in [(1.0, 1.0), (1.5, -1.0), (11 / 6, 1.0), (11 / 6, -1.0)]:
        got = wlc.test_solve(L, k, cd, sigma, 0.004, 0.004 * 0.9, v)
        for p in range(64):
            ref = oracle.ginv(L, cd, sigma, 0.004, 0.9, k[p]) @ v[p]
            assert _rel(got[p], ref) < 1e-13


CASES = [
    # (shape (nz, ny, nx), box, Nblk, Lmax, lambda, force generic z kernel)
    ((4, 6, 8), (4.0, 2.5, 3.0), (4, 5, 3), (2, 3, 1), (1.0, 0.5, 2.0), 0),
    ((6, 12, 6), (2.0, 3.0, 2.5), (3, 5, 4), (1, 3, 2), (2.0, 0.3, 1.0), 0),
    ((64, 6, 8), (3.0, 2.0, 6.0), (5, 4, 6), (3, 2, 4), (0.7, 1.3, 0.4), 0),
    ((64, 6, 8), (3.0, 2.0, 6.0), (5, 4, 6), (3, 2, 4), (0.7, 1.3, 0.4), 1),
    ((8, 8, 8), (4.0, 4.0, 4.0), (9,), (4,), (0.7,), 0),
    ((24, 4, 6), (2.0, 2.0, 5.0), (6, 7), (0, 2), (1.0, 0.0), 0),
    ((2, 2, 2), (1.0, 1.0, 1.0), (2, 1, 2), (8, 5, 6), (0.5, 1.0, 3.0), 0),
    ((16, 4, 4), (2.0, 2.0, 2.0), (3, 3, 3), (6, 6, 6), (1.0, 1.0, 1.0), 0),
    # nz = 64 at Lmax 8: the one-wave kernel with all 81 harmonics of a column
    ((64, 4, 4), (3.0, 2.0, 2.0), (3, 4), (8, 4), (0.5, 1.0), 0),
    # nz = 64 with Lmax 5..7: the split z-solve kernel (both parts, crossing chains), one
    # 4-kx tile per workgroup
    ((64, 4, 6), (3.0, 2.0, 2.5), (4, 3, 4), (6, 5, 7), (1.0, 0.7, 1.3), 0),
    # ... with nx/2 + 1 = 5 columns: a partial last tile (one real column of four)
    ((64, 6, 8), (3.0, 2.0, 4.0), (3, 4, 3), (7, 6, 5), (0.9, 1.1, 0.6), 0),
    # 64 x 64 planes (register-resident plane kernels), basis shrinks then grows
    ((4, 64, 64), (2.0, 5.0, 4.0), (3, 4, 3), (3, 2, 3), (1.0, 0.5, 2.0), 0),
    ((2, 64, 64), (1.0, 3.0, 3.0), (5,), (4,), (0.8,), 0),
    # nz = 128 (two waves per column, radix-2 step across them) at Lmax 8 and 7, and the
    # generic z kernel on the same case
    ((128, 4, 6), (6.0, 2.0, 2.5), (3, 4), (8, 7), (0.5, 1.0), 0),
    ((128, 4, 6), (6.0, 2.0, 2.5), (3, 4), (8, 7), (0.5, 1.0), 1),
    # ... nx/2 + 1 = 5 columns (a half-padded and an all-padding workgroup), Lmax 5 and 3
    ((128, 4, 8), (5.0, 2.0, 3.0), (4, 3), (5, 3), (1.0, 0.4), 0),
    # 24 x 24 and 48 x 48 planes (compile-time Stockham passes: radix 8, 3 and 4, 4, 3)
    ((4, 24, 24), (2.0, 4.0, 4.0), (3, 4, 3), (4, 3, 5), (1.0, 0.6, 1.4), 0),
    ((2, 48, 48), (1.5, 5.0, 5.0), (4, 3), (3, 4), (0.8, 1.2), 0),
    # 128 x 128 planes (128 KB of LDS per workgroup, radix 16 x 8 passes)
    ((4, 128, 128), (2.0, 6.0, 6.0), (3, 3), (2, 1), (0.5, 1.5), 0),
    ((2, 128, 128), (1.0, 4.0, 5.0), (2, 3, 2), (3, 2, 3), (1.0, 0.5, 2.0), 0),
    # ... nz = 16 (the XCD z-range block order with two z-planes per XCD)
    ((16, 128, 128), (2.0, 6.0, 6.0), (3, 3), (2, 1), (0.5, 1.5), 0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[2]}-L{c[3]}-g{c[5]}" for c in CASES])
def test_sweep_matches_oracle(wlc, oracle, case, monkeypatch):
    shape, box, Nblk, L, lam, generic = case
    monkeypatch.setenv("WLC_GENERIC_ZSOLVE", str(generic))
    rng = np.random.default_rng(sum(shape) + len(Nblk))
    w = 1.5 * rng.standard_normal((len(Nblk),) + shape)
    orc = oracle.Oracle(shape, box, Nblk, L, lam)
    q = orc.forward(w)
    qs = orc.backward(w)
    phi, Q, td = orc.density(q, qs)
    g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, keep_qstar=True)
    g.set_w(w)
    g.MDE_q()
    g.MDE_qstar()
    gphi, gQ, gtd = g.density()
    for s in range(orc.Ns + 1):
        assert _rel(g.get_q(+1, s), orc.slice(q, s)) < TOL, f"forward slice {s}"
        assert _rel(g.get_q(-1, s), orc.slice(qs, s)) < TOL, f"backward slice {s}"
    assert abs(gQ - Q) / abs(Q) < TOL
    assert abs(gtd - td) / abs(td) < TOL
    assert _rel(gphi, phi) < TOL
    g.close()


@pytest.mark.parametrize("name", ["ab_even", "abc_odd_radix3", "homopolymer"])
def test_sweep_matches_golden_fixture(wlc, name):
    f = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    g = wlc.WormlikeMDE(tuple(f["shape"]), tuple(f["box"]), tuple(f["Nblk"]), tuple(f["L"]),
                        tuple(f["lam"]), keep_qstar=True)
    g.set_w(f["w"])
    g.MDE_q()
    g.MDE_qstar()
    phi, Q, td = g.density()
    Ns = int(sum(f["Nblk"]))
    assert _rel(g.get_q(+1, Ns), f["q_end"]) < TOL
    assert _rel(g.get_q(+1, int(f["s_mid"])), f["q_mid"]) < TOL
    assert _rel(g.get_q(-1, 0), f["qs_0"]) < TOL
    assert _rel(phi, f["phi"]) < TOL
    assert abs(Q - float(f["Q"])) / float(f["Q"]) < TOL
    g.close()


def test_state_errors(wlc):
    g = wlc.WormlikeMDE((4, 4, 4), (2.0, 2.0, 2.0), (3, 3), (2, 2), (1.0, 1.0))
    with pytest.raises(wlc.WlcError) as e:
        g.MDE_q()
    assert e.value.code == -4
    g.set_w(np.zeros((2, 4, 4, 4)))
    with pytest.raises(wlc.WlcError) as e:
        g.density()
    assert e.value.code == -4
    g.MDE_q()
    g.MDE_qstar()
    g.density()
    with pytest.raises(wlc.WlcError) as e:
        g.get_q(-1, 3)               # backward slices need keep_qstar
    assert e.value.code == -4
    g.get_q(-1, 0)
    with pytest.raises(ValueError):
        g.set_w(np.zeros((3, 4, 4, 4)))
    g.set_w(np.zeros((2, 4, 4, 4)))  # a new w invalidates the sweeps
    with pytest.raises(wlc.WlcError):
        g.density()
    g.close()


def test_set_box_changes_the_solution(wlc, oracle):
    shape, Nblk, L, lam = (8, 8, 8), (4, 4), (3, 3), (1.0, 1.0)
    rng = np.random.default_rng(9)
    w = rng.standard_normal((2,) + shape)
    g = wlc.WormlikeMDE(shape, (3.0, 3.0, 3.0), Nblk, L, lam)
    g.set_w(w)
    g.set_box((4.0, 3.5, 5.0))
    g.set_w(w)
    g.MDE_q()
    g.MDE_qstar()
    phi, Q, _ = g.density()
    orc = oracle.Oracle(shape, (4.0, 3.5, 5.0), Nblk, L, lam)
    ophi, oQ, _ = orc.density(orc.forward(w), orc.backward(w))
    assert abs(Q - oQ) / oQ < TOL
    assert _rel(phi, ophi) < TOL
    g.close()


FUSE_CASES = [
    # compile-time 32x32 planes, basis grows then shrinks at the junctions
    ((8, 32, 32), (3.0, 3.0, 2.0), (5, 6, 4), (4, 6, 3), (1.0, 0.5, 2.0)),
    # compile-time 64x64 planes, shrinking then growing basis, generic z kernel
    ((6, 64, 64), (5.0, 5.0, 2.0), (4, 3, 5), (6, 2, 5), (1.0, 1.0, 1.0)),
    # runtime plane sizes
    ((12, 6, 18), (2.0, 2.5, 3.0), (3, 4, 4), (2, 3, 1), (0.8, 1.2, 0.6)),
    # 128x128 planes (dynamic 128 KB LDS, 512 threads), basis grows then shrinks
    ((2, 128, 128), (2.0, 6.0, 5.0), (3, 3, 3), (1, 3, 2), (0.7, 1.0, 1.4)),
]


@pytest.mark.parametrize("case", FUSE_CASES, ids=[f"{c[0]}-L{c[3]}" for c in FUSE_CASES])
def test_fused_inverse_rhs_is_bit_identical(wlc, case, monkeypatch):
    """The inverse xy pass of step s fused with the right-hand side of step s+1 keeps
    q(s) on chip instead of reloading it: every slice must equal the unfused sweep's."""
    shape, box, Nblk, L, lam = case
    rng = np.random.default_rng(77)
    w = 2.0 * rng.standard_normal((len(Nblk),) + shape)
    res = []
    for fuse in ("0", "1"):
        monkeypatch.setenv("WLC_FUSE", fuse)
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, keep_qstar=True)
        g.set_w(w)
        g.MDE_q()
        g.MDE_qstar()
        phi, Q, td = g.density()
        Ns = sum(Nblk)
        res.append(([g.get_q(+1, s) for s in range(Ns + 1)],
                    [g.get_q(-1, s) for s in range(Ns + 1)], phi, Q, td))
        g.close()
    (q0, qs0, phi0, Q0, td0), (q1, qs1, phi1, Q1, td1) = res
    for s in range(len(q0)):
        assert np.array_equal(q0[s], q1[s]), f"forward slice {s}"
        assert np.array_equal(qs0[s], qs1[s]), f"backward slice {s}"
    assert np.array_equal(phi0, phi1) and Q0 == Q1 and td0 == td1


@pytest.mark.parametrize("shape,L", [((64, 6, 8), (5, 7, 6)), ((128, 4, 8), (5, 7, 6)),
                                     ((128, 4, 6), (8, 6, 8)), ((48, 6, 8), (6, 4, 7)),
                                     ((48, 4, 6), (5, 7, 4)),
                                     ((24, 6, 12), (4, 5, 7)), ((32, 4, 6), (6, 4, 4))])
def test_split_z_solve_matches_one_wave_kernel(wlc, monkeypatch, shape, L):
    """Lmax 5..7 at nz = 64 and 5..8 at nz = 128 runs the z solve with the harmonics of a
    column split over two waves (per z half at nz = 128); the unsplit kernel computes the same step
    (equal up to rounding: the split kernel passes k = 0 through identity-angle rotations
    instead of a diagonal shortcut).  nx/2 + 1 = 5 leaves padding columns in the last tile.
    At nz = 64 the whole-tile and half-tile split kernels and the two-column split kernel
    run the same operations (equal up to the compiler's multiply-add contraction); at nz = 24,
    32 and 48 the half-tile kernel (wave-local 3, 4 or 6 x 8 transform) is compared with the
    generic split kernel (block-wide Stockham transform).  At
    other nz (24: the two parts share a wave; 32) the generic kernel splits Lmax 4..7 over
    two threads per k-point."""
    box, Nblk, lam = (3.0, 2.0, 5.0), (3, 4, 3), (1.2, 0.6, 1.0)
    rng = np.random.default_rng(31)
    w = 1.5 * rng.standard_normal((3,) + shape)
    out = []
    monkeypatch.setenv("WLC_GENERIC_ZSOLVE", "0")
    splits = {64: ("0", "1", "2", "3")}.get(shape[0], ("0", "1", "3") if shape[0] in (24, 32, 48) else ("0", "1"))
    for split in splits:
        monkeypatch.setenv("WLC_ZSOLVE_SPLIT", split)
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, keep_qstar=True)
        g.set_w(w)
        g.MDE_q()
        g.MDE_qstar()
        phi, Q, _ = g.density()
        out.append((g.get_q(+1, sum(Nblk)), g.get_q(-1, 0), phi, Q))
        g.close()
    (q0, s0, p0, Q0), (q1, s1, p1, Q1) = out[:2]
    assert _rel(q1, q0) < 1e-13 and _rel(s1, s0) < 1e-13 and _rel(p1, p0) < 1e-13
    assert abs(Q1 - Q0) / abs(Q0) < 1e-13
    names = {"2": "tile", "3": "half-tile"}
    for k in range(2, len(out)):
        q2, s2, p2, Q2 = out[k]
        print(names[splits[k]], "vs split kernel 1:", _rel(q2, q1), _rel(s2, s1), _rel(p2, p1), flush=True)
        assert _rel(q2, q1) < 1e-13 and _rel(s2, s1) < 1e-13 and _rel(p2, p1) < 1e-13
        assert abs(Q2 - Q1) / abs(Q1) < 1e-13


@pytest.mark.parametrize("keep", [False, True])
def test_concurrent_sweeps_match_sequential(wlc, monkeypatch, keep):
    """When the full backward store fits, the backward sweep runs concurrently with the
    forward sweep (own stream and spectrum) and each density term runs once both slices
    exist; the sequential path (ring of four backward slices) must give the same result.
    Propagator slices are bit-identical; the density sums in a different order."""
    shape, box, Nblk, L, lam = (16, 32, 32), (2.0, 4.0, 4.0), (5, 6, 4), (3, 4, 2), (1.0, 0.6, 1.5)
    rng = np.random.default_rng(12)
    w = 1.2 * rng.standard_normal((3,) + shape)
    res = []
    for conc in ("0", "1"):
        monkeypatch.setenv("WLC_CONCURRENT", conc)
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, keep_qstar=keep)
        g.set_w(w)
        for _ in range(2):  # two iterations: buffers and events are reused
            g.MDE_q()
            g.MDE_qstar()
            phi, Q, td = g.density()
        res.append((phi, Q, td, g.get_q(+1, sum(Nblk)), g.get_q(-1, 0)))
        g.close()
    (p0, Q0, t0, q0, s0), (p1, Q1, t1, q1, s1) = res
    assert np.array_equal(q0, q1) and np.array_equal(s0, s1)
    assert Q0 == Q1
    assert abs(t1 - t0) / abs(t0) < 1e-13 and _rel(p1, p0) < 1e-13


@pytest.mark.parametrize("K,stored", [(3, 0), (5, 0), (8, 0), (40, 0), (3, 2), (5, 1), (8, 2)])
def test_checkpointed_forward_store_matches_full_store(wlc, monkeypatch, K, stored):
    """With ckpt_every = K the forward sweep keeps only segment-boundary checkpoints and the
    backward sweep recomputes each segment (on its own stream, double-buffered) before its
    densities; with `stored` > 0 the lowest segments are kept whole (the hybrid store of the
    automatic mode) and only the ones above are recomputed.  phi, Q, totden, q(Ns), q*(0)
    and the stored forward slices are bit-identical to the full store."""
    monkeypatch.setenv("WLC_CONCURRENT", "0")
    monkeypatch.setenv("WLC_STORED_SEGMENTS", str(stored))
    shape, box, Nblk, L, lam = (8, 16, 16), (2.0, 3.0, 3.0), (7, 9, 7), (3, 4, 2), (1.0, 0.6, 1.5)
    rng = np.random.default_rng(21)
    w = 1.5 * rng.standard_normal((3,) + shape)
    res = []
    top = stored * K  # last stored forward slice
    for k in (0, K):
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, ckpt_every=k)
        g.set_w(w)
        for _ in range(2):
            g.MDE_q()
            g.MDE_qstar()
            phi, Q, td = g.density()
        res.append((phi, Q, td, g.get_q(+1, sum(Nblk)), g.get_q(-1, 0),
                    [g.get_q(+1, s) for s in range(top + 1)] if stored else []))
        if k:
            with pytest.raises(wlc.WlcError) as e:
                g.get_q(+1, top + 1)
            assert e.value.code == -4
        g.close()
    (p0, Q0, t0, q0, s0, f0), (p1, Q1, t1, q1, s1, f1) = res
    assert np.array_equal(q0, q1) and np.array_equal(s0, s1)
    assert Q0 == Q1 and t0 == t1 and np.array_equal(p0, p1)
    assert all(np.array_equal(a, b) for a, b in zip(f0, f1))


@pytest.mark.parametrize("mode", [1, 2])
def test_timing_counts_launches_without_changing_results(wlc, mode):
    """Per-kernel timing brackets every launch with events on its own stream (mode 1: read
    back at the density call; mode 2: also finishing the forward sweep before the backward
    one): the launch counts follow the sweep structure and the results are bitwise those of
    an untimed run."""
    shape, box, Nblk, L, lam = (8, 16, 16), (2.0, 3.0, 3.0), (4, 5, 3), (2, 3, 1), (1.0, 0.6, 1.5)
    rng = np.random.default_rng(41)
    w = rng.standard_normal((3,) + shape)
    g = wlc.WormlikeMDE(shape, box, Nblk, L, lam)
    g.set_w(w)
    g.MDE_q()
    g.MDE_qstar()
    phi0, Q0, td0 = g.density()
    g.set_timing(mode)
    for _ in range(2):
        g.MDE_q()
        g.MDE_qstar()
        phi1, Q1, td1 = g.density()
    ms, cnt = g.get_timing()
    g.set_timing(0)
    with pytest.raises(wlc.WlcError):
        g.set_timing(3)
    g.close()
    assert np.array_equal(phi1, phi0) and Q1 == Q0 and td1 == td0
    Ns = sum(Nblk)
    # per sweep pair: Ns z solves and one A, Ns - 1 fused C+A, one C per direction;
    # Ns + 1 density terms
    assert cnt[1] == 2 * 2 * Ns
    assert cnt[0] == 2 * 2 and cnt[2] == 2 * 2 and cnt[4] == 2 * 2 * (Ns - 1)
    assert cnt[3] == 2 * (Ns + 1)
    assert all(m > 0.0 for m in ms)


# FP32 half spectrum (all arithmetic FP64, the spectrum rounded to FP32 where it is stored,
# twice per contour step): stated tolerance against the FP64 oracle 2e-6 relative in phi,
# Q and q(Ns) on these short chains (observed 1e-8 .. 3e-7).
FP32_TOL = 2e-6
FP32_CASES = [
    ((64, 4, 6), (3.0, 2.0, 2.5), (4, 3, 4), (6, 5, 7), (1.0, 0.7, 1.3)),     # split z, nz 64
    ((4, 64, 64), (2.0, 5.0, 4.0), (3, 4, 3), (3, 2, 3), (1.0, 0.5, 2.0)),    # 64x64 planes
    ((2, 128, 128), (1.0, 4.0, 5.0), (2, 3, 2), (3, 2, 3), (1.0, 0.5, 2.0)),  # 128x128 planes
    ((128, 4, 6), (6.0, 2.0, 2.5), (3, 4), (8, 7), (0.5, 1.0)),               # nz 128, both kernels
    ((6, 12, 6), (2.0, 3.0, 2.5), (3, 5, 4), (1, 3, 2), (2.0, 0.3, 1.0)),     # generic kernels
    ((8, 8, 8), (4.0, 4.0, 4.0), (9,), (4,), (0.7,)),                         # one-wave z, nz 8
]


@pytest.mark.parametrize("case", FP32_CASES, ids=[f"{c[0]}-L{c[3]}" for c in FP32_CASES])
def test_fp32_spectrum_matches_oracle_within_stated_tolerance(wlc, oracle, case):
    shape, box, Nblk, L, lam = case
    rng = np.random.default_rng(sum(shape))
    w = 1.5 * rng.standard_normal((len(Nblk),) + shape)
    orc = oracle.Oracle(shape, box, Nblk, L, lam)
    q = orc.forward(w)
    qs = orc.backward(w)
    phi, Q, td = orc.density(q, qs)
    out = []
    for fp32 in (False, True):
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, spectrum_fp32=fp32)
        g.set_w(w)
        g.MDE_q()
        g.MDE_qstar()
        gphi, gQ, _ = g.density()
        out.append((gphi, gQ, g.get_q(+1, sum(Nblk))))
        g.close()
    (p64, Q64, q64), (p32, Q32, q32) = out
    assert _rel(p32, phi) < FP32_TOL and abs(Q32 - Q) / Q < FP32_TOL
    assert _rel(q32, orc.slice(q, sum(Nblk))) < FP32_TOL
    # the mode really stores FP32: the result differs from the FP64 path
    assert _rel(p32, p64) > 1e-12


def test_spectrum_precision_switch_between_sweeps(wlc):
    """wlc_mde_set_spectrum_fp32 switches the storage precision between sweeps: the results
    equal, bit for bit, those of a handle created in that mode, and sweeps from before the
    switch are not reused."""
    shape, box, Nblk, L, lam = (8, 16, 16), (2.0, 3.0, 3.0), (4, 5, 3), (2, 3, 1), (1.0, 0.6, 1.5)
    rng = np.random.default_rng(43)
    w = rng.standard_normal((3,) + shape)

    def run(g):
        g.set_w(w)
        g.MDE_q()
        g.MDE_qstar()
        return g.density()

    ref = {}
    for fp32 in (False, True):
        g = wlc.WormlikeMDE(shape, box, Nblk, L, lam, spectrum_fp32=fp32)
        ref[fp32] = run(g)
        g.close()
    g = wlc.WormlikeMDE(shape, box, Nblk, L, lam)
    run(g)
    g.set_spectrum_fp32(True)
    with pytest.raises(wlc.WlcError):
        g.density()
    for fp32 in (True, False, True):
        g.set_spectrum_fp32(fp32)
        phi, Q, td = run(g)
        assert np.array_equal(phi, ref[fp32][0]) and Q == ref[fp32][1] and td == ref[fp32][2]
    g.close()
